"""Benchmark: atom E+F evaluations per second on one MI355X.

Workload (one "step"): one FP32 energy + forces + virial evaluation of Cu fcc 20x20x20 (32,000 atoms,
a = 3.61 A, Gaussian jitter sigma = 0.05 A, seed 1) with the synthetic 51-64-64-1 tanh potential
(maise_amd/potentials/cu_51_64_64.nn: 8 G2 + 43 G4, Rc = 6.0 A).  The structure is uploaded once; the timed
region covers neighbour build, descriptors, MLP forward/backward, forces/virial and the FP64 reductions of
K evaluations, inputs resident in HBM.

The JSON line also carries
  roofline      the dominant kernel (the descriptor or the force/virial kernel, FP32 vector-ALU bound) priced against
                the MI355X FP32 peak: algorithmic flops per launch (exact neighbour and triplet counts of the
                structure) / mean kernel time measured with HIP events on the library's stream over the timed region;
                beside it the roofline model of the whole evaluation (T_roofline / T_measured with neighbour bytes
                against HBM, descriptor+force flops and transcendentals against the FP32 peak and the transcendental
                rate measured on the device by the library's microbenchmark, MLP flops against the FP32 MFMA peak)
                and each stage's fraction of its own bound;
  cpu_baseline  the reference's own CPU path (compiled from its sources, oracle/_ref/maise_ref) timed on bounded
                samples of the same workload: Cu fcc 10x10x10 (4,000 atoms) on the host's cores (at most the 16-core
                share of one GPU) and Cu fcc 4x4x4 (256 atoms) on one core, with nproc and the CPU model;
  error_fp64    max |dE|/atom and |dF| of the FP32 path against the FP64 oracle on the 256-atom cell.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

POT = os.path.join(ROOT, "maise_amd", "potentials", "cu_51_64_64.nn")
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def algorithmic_counts(c, lat, rc):
    """Exact sum over atoms of M (neighbours) and K (neighbour pairs j<k with r_jk < rc) from the GPU list."""
    row, j, img = c.neighbors()
    X = c.result(0, forces=False)["X"]
    N = len(row) - 1
    M = int(row[-1])
    K = 0
    d_all = X[j] + img @ lat - np.repeat(X[:N], np.diff(row), axis=0)
    counts = np.diff(row)
    rc2 = rc * rc
    o = 0
    # group atoms with equal row length so the pair distances vectorise
    for n in np.unique(counts):
        idx = np.nonzero(counts == n)[0]
        for s in range(0, len(idx), 2048):
            sel = idx[s:s + 2048]
            starts = row[sel]
            d = d_all[starts[:, None] + np.arange(n)[None, :]]          # [B, n, 3]
            diff = d[:, :, None, :] - d[:, None, :, :]
            r2 = np.einsum("bjkq,bjkq->bjk", diff, diff)
            iu = np.triu_indices(n, 1)
            K += int(np.count_nonzero(r2[:, iu[0], iu[1]] < rc2))
        o += len(idx)
    return M, K


def flops_model(M, K, ng2=8, neta=7, ng4=43):
    """Algorithmic FP32 flops of the descriptor kernel and of the force/virial kernel (per launch)."""
    desc = M * (11 + 3 * ng2) + K * (48 + 2 * neta + 2 * ng4)
    force = M * (30 + 8 * ng2) + K * (2 * (48 + 2 * neta + 2 * ng4) + 6 * ng4 + 60)
    return desc, force


def host_cpu_info():
    """(logical CPUs this process may run on, CPU model name) of the host."""
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return nproc, model


def progress(msg):
    """One line per finished step of the long legs, on stderr (the result stays the one JSON line on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(threads):
    """The reference's own CPU path (CALL_MAISE compiled from its sources, oracle/_ref) on bounded samples of the same
    lattice and potential: Cu fcc 10x10x10 (4,000 atoms) on `threads` OpenMP threads, and Cu fcc 4x4x4 (256 atoms) on
    one thread.  Each figure is the mean of 2 calls after one warm-up, model load excluded."""
    from maise_amd import workloads as wl
    from oracle import pyoracle as po
    nproc, cpu_model = host_cpu_info()
    info = {"nproc": nproc, "cpu_model": cpu_model}
    if po.have_reference():
        s, l, x = wl.cu_fcc(10, 3.61, 0.05, seed=1)
        r = po.run_reference(POT, s, l, x, nm=200, repeats=2, threads=threads, timeout=900)
        s1, l1, x1 = wl.cu_fcc(4, 3.61, 0.05, seed=1)
        r1 = po.run_reference(POT, s1, l1, x1, nm=200, repeats=2, threads=1, timeout=900)
        out = {"value": len(s) / r["time"], "unit": "atom E+F evals/s", "cores": threads, "kind": "reference",
               "sample": f"reference CALL_MAISE (compiled from its sources, OpenMP {threads} threads) on Cu fcc "
                         f"10x10x10 = {len(s)} atoms, same potential and lattice; mean of 2 calls after one warm-up "
                         "(model load excluded)",
               "single_core": {"value": len(s1) / r1["time"], "unit": "atom E+F evals/s", "cores": 1,
                               "sample": f"the same on Cu fcc 4x4x4 = {len(s1)} atoms, one thread"}}
        out.update(info)
        return out
    m = po.OracleModel(POT)
    s, l, x = wl.cu_fcc(6, 3.61, 0.05, seed=1)
    m.evaluate(s, l, x)
    t0 = time.perf_counter()
    m.evaluate(s, l, x)
    t = time.perf_counter() - t0
    out = {"value": len(s) / t, "unit": "atom E+F evals/s", "cores": po.num_threads(), "kind": "port",
           "sample": f"FP64 oracle on Cu fcc 6x6x6 = {len(s)} atoms"}
    out.update(info)
    return out


def leg_cpu_baselines(threads, batch):
    """CPU figures beside the legs, on bounded samples: MD = the reference's CALL_MAISE (compiled from its sources)
    on the Cu-Ag alloy at 8x8x8 (2,048 atoms, same lattice and potential; an MD step costs one E+F evaluation), mean
    of 2 calls after a warm-up; batched FIRE and the search = FIRE on FP64 oracle forces (the reference relaxes with
    GSL's BFGS, absent here), one structure after another, on the first 2 structures of the Cu-Pd-Ag batch and on 4
    random Mg16Ca8 cells of the search's initial population."""
    from maise_amd import workloads as wl
    from maise_amd.evolve import GAConfig, random_structure
    from oracle import pyoracle as po
    out = {}
    if po.have_reference():
        s, l, x = wl.alloy_fcc(8, 3.8, 2, 0.0, seed=2)
        r = po.run_reference(MD_POT, s, l, x, nm=300, repeats=2, threads=threads, timeout=900)
        out["md"] = {"value": len(s) / r["time"], "unit": "atom-steps/s", "cores": threads, "kind": "reference",
                     "sample": f"reference CALL_MAISE, Cu-Ag 50/50 fcc 8x8x8 ({len(s)} atoms), one E+F evaluation "
                               "per MD step, mean of 2 calls after a warm-up"}
    if batch is not None:
        m = po.OracleModel(RELAX_POT)
        t0 = time.perf_counter()
        for s, l, x in batch[:2]:
            po.fire_relax(m, s, l, x, urep=100.0)
        out["relax"] = {"value": 2 / (time.perf_counter() - t0), "unit": "relaxed structures/s",
                        "cores": po.num_threads(), "kind": "port",
                        "sample": "FIRE on FP64 oracle forces, the first 2 structures of the Cu-Pd-Ag batch"}
    cfg = GAConfig()
    rng = np.random.default_rng(cfg.seed)
    cells = [random_structure(cfg, rng) for _ in range(4)]
    m = po.OracleModel(GA_POT)
    t0 = time.perf_counter()
    for s, l, x in cells:
        po.fire_relax(m, s, l, x, urep=cfg.urep)
    out["search"] = {"value": 4 / (time.perf_counter() - t0), "unit": "relaxed offspring/s",
                     "cores": po.num_threads(), "kind": "port",
                     "sample": "FIRE on FP64 oracle forces, 4 random Mg16Ca8 cells of the initial population"}
    return out


def measured_rates(maise_amd):
    """Issue rates measured on the device by the library's microbenchmark (lane operations per second)."""
    names = ("fma", "exp2", "cos", "sqrt", "rcp")
    return {n: maise_amd.calibrate_rate(k) for k, n in enumerate(names)}


def roofline_model(N, M, K, stages, ms, model, rates, ng2=8, neta=7, ng4=43):
    """Roofline model v1 of the evaluation: T_roofline = neighbour bytes / HBM bandwidth + max(descriptor and force
    flops / FP32 peak, their transcendentals / the measured transcendental rate) + MLP flops / FP32 MFMA peak, and the
    fraction of each stage's own bound reached at its measured time."""
    D, H1, H2 = model.D, model.H1, model.H2
    nbr_bytes = N * (28 + 4) + 8 * M
    fl_desc, fl_force = flops_model(M, K, ng2, neta, ng4)
    fl_df = fl_desc + fl_force
    trans = M * (5 + 2 * ng2) + K * (12 + 3 * neta)
    fl_mlp = N * (2 * (D * H1 + H1 * H2 + H2) + 2 * (H2 * H1 + H1 * D) + 2 * H2)
    p_trans = rates["exp2"]
    t_nbr = nbr_bytes / (HBM_PEAK_GBS * 1e9)
    t_df = max(fl_df / (FP32_PEAK_TFLOPS * 1e12), trans / p_trans)
    t_mlp = fl_mlp / (FP32_PEAK_TFLOPS * 1e12)
    t_roof = t_nbr + t_df + t_mlp
    s_nbr, s_desc, s_mlp, s_force = (stages[0] * 1e-3, stages[1] * 1e-3, stages[2] * 1e-3, stages[3] * 1e-3)
    return {
        "t_roofline_ms": round(t_roof * 1e3, 4),
        "t_measured_ms": round(ms, 4),
        "t_roofline_over_t_measured": round(t_roof * 1e3 / ms, 4),
        "per_stage": {
            "neighbours": {"bound": "hbm", "bytes": nbr_bytes, "achieved_GBs": round(nbr_bytes / s_nbr / 1e9, 1),
                           "peak_GBs": HBM_PEAK_GBS, "frac": round(nbr_bytes / s_nbr / 1e9 / HBM_PEAK_GBS, 4)},
            "descriptors_forces": {"bound": "valu", "flops": fl_df, "transcendentals": trans,
                                   "achieved_TFLOPs": round(fl_df / (s_desc + s_force) / 1e12, 2),
                                   "frac_fp32_peak": round(fl_df / (s_desc + s_force) / 1e12 / FP32_PEAK_TFLOPS, 4),
                                   "achieved_trans_per_s": trans / (s_desc + s_force),
                                   "frac_trans_rate": round(trans / (s_desc + s_force) / p_trans, 4)},
            "mlp": {"bound": "mfma", "flops": fl_mlp, "achieved_TFLOPs": round(fl_mlp / s_mlp / 1e12, 2),
                    "peak_TFLOPs": FP32_PEAK_TFLOPS, "frac": round(fl_mlp / s_mlp / 1e12 / FP32_PEAK_TFLOPS, 4)},
        },
        "peaks": {"hbm_GBs": HBM_PEAK_GBS, "fp32_TFLOPs": FP32_PEAK_TFLOPS, "fp32_mfma_TFLOPs": FP32_PEAK_TFLOPS,
                  "trans_per_s_measured": p_trans},
        "measured_rates_per_s": {k: float(v) for k, v in rates.items()},
    }


RELAX_POT = os.path.join(ROOT, "tests", "golden", "models", "nn_CuPdAg_d3_v1")
MD_POT = os.path.join(ROOT, "tests", "golden", "models", "nn_CuAg_d3_v1")


def relax_leg(maise_amd, batch, args):
    """Batched FIRE of the Cu-Pd-Ag batch (one context, every structure at once, converged ones compacted away)
    against the same FIRE run one structure per context (the plain path), to fmax 0.01 eV/A."""
    model = maise_amd.Model(RELAX_POT)
    max_steps, fmax = 3000, 0.01
    c = maise_amd.Context(model, device=0, urep=100.0)
    c.set_batch(batch[:64])
    c.relax(max_steps=50, fmax=fmax)           # warm-up: code objects, allocations
    c.set_autotune(True)                       # tuned on the full batch before the first capture (timed)
    c.set_skin(args.relax_skin)                # skin neighbour lists (results bit-identical to rebuilding each step)
    t0 = time.perf_counter()
    c.set_batch(batch)
    steps, done = c.relax(max_steps=max_steps, fmax=fmax)
    dt = time.perf_counter() - t0
    fres = c.batch_result()
    shape = c.shape()
    log = c.relax_log()
    c.close()
    # wall time by the share of structures still active in each segment between host checks
    S0 = float(len(batch))

    def wall_split(log):
        split = {}
        for name, lo, hi in (("active_ge_50pct", 0.5, 2.0), ("active_10_50pct", 0.1, 0.5),
                             ("active_lt_10pct", 0.0, 0.1)):
            sel = (log[:, 1] / S0 >= lo) & (log[:, 1] / S0 < hi)
            split[name] = {"seconds": float(log[sel, 3].sum()),
                           "iterations": int(np.diff(np.concatenate([[0], log[:, 0]]))[sel].sum())}
        return split
    split = wall_split(log)
    # the same batch with the device L-BFGS (the reference minimises with GSL's BFGS2; no line search here)
    c = maise_amd.Context(model, device=0, urep=100.0)
    c.set_batch(batch[:64])
    c.relax_lbfgs(max_steps=50, fmax=fmax)     # warm-up
    c.set_autotune(True)
    c.set_skin(args.relax_skin)
    t0 = time.perf_counter()
    c.set_batch(batch)
    lsteps, ldone = c.relax_lbfgs(max_steps=max_steps, fmax=fmax)
    ldt = time.perf_counter() - t0
    lres = c.batch_result()
    lsplit = wall_split(c.relax_log())
    c.close()
    nat = np.array([len(b[0]) for b in batch], dtype=np.float64)
    ef, el = fres["E"] / nat, lres["E"] / nat
    lbfgs = {"relaxed_structures_per_s": float(ldone.sum()) / ldt, "seconds": ldt, "converged": int(ldone.sum()),
             "steps_mean": float(lsteps.mean()), "steps_max": int(lsteps.max()), "memory": 20,
             "speedup_over_fire": (float(ldone.sum()) / ldt) / (float(done.sum()) / dt),
             "relaxed_energy_per_atom_mean_eV": {"fire": float(ef.mean()), "lbfgs": float(el.mean())},
             "share_lower_than_fire_by_1meV": float(np.mean(el < ef - 1e-3)),
             "share_higher_than_fire_by_1meV": float(np.mean(el > ef + 1e-3)),
             "wall_by_active_share": lsplit,
             "tail_share_of_wall": lsplit["active_lt_10pct"]["seconds"] / max(ldt, 1e-12)}
    progress(f"relax: L-BFGS {lbfgs['relaxed_structures_per_s']:.0f} structures/s, steps mean "
             f"{lbfgs['steps_mean']:.1f} max {lbfgs['steps_max']}")
    nsingle = min(args.relax_single, len(batch))
    c1 = maise_amd.Context(model, device=0, urep=100.0)
    t1 = time.perf_counter()
    s1 = []
    for k in range(nsingle):
        c1.set_batch([batch[k]])
        st, _ = c1.relax(max_steps=max_steps, fmax=fmax)
        s1.append(int(st[0]))
    dt1 = time.perf_counter() - t1
    c1.close()
    return {"workload": f"{len(batch)} Cu-Pd-Ag structures, 8-64 atoms ({sum(len(b[0]) for b in batch)} atoms), "
                        "random triclinic cells, 15-20 A^3/atom, dmin 2.0 A, seed 3; shipped potential "
                        "nn_CuPdAg_d3_v1 (282-10-10-1, Rc 7.5 A), UREP 100 eV; FIRE to fmax 0.01 eV/A, cap 3000",
            "relaxed_structures_per_s": float(done.sum()) / dt, "seconds": dt,
            "converged": int(done.sum()), "structures": len(batch), "skin_A": args.relax_skin,
            "fire_steps_mean": float(steps.mean()), "fire_steps_max": int(steps.max()),
            "autotuned_shape": {"nbr_wpb": shape[0], "desc_wpb": shape[1], "force_wpb": shape[2],
                                "pair_store": shape[3]},
            "wall_by_active_share": split,
            "tail_share_of_wall": split["active_lt_10pct"]["seconds"] / max(float(log[:, 3].sum()), 1e-12),
            "single_structure_path": {"structures": nsingle, "seconds": dt1,
                                      "relaxed_structures_per_s": nsingle / dt1,
                                      "steps_equal_batched": bool(np.array_equal(np.array(s1), steps[:nsingle]))},
            "batched_over_single": (float(done.sum()) / dt) / (nsingle / dt1),
            "lbfgs": lbfgs}


def md_leg(maise_amd, args):
    """NVT molecular dynamics of the Cu-Ag 50/50 fcc 12x12x12 alloy (6912 atoms, a = 3.8 A), 600 K
    Maxwell-Boltzmann start (seed 1), 1 fs steps, state on the device; with skin neighbour lists (1 A; the
    trajectory is bit-identical to rebuilding the lists every step, which is timed too)."""
    from maise_amd import workloads as wl
    s, l, x = wl.alloy_fcc(12, 3.8, 2, 0.0, seed=2)
    model = maise_amd.Model(MD_POT)

    def run(skin):
        c = maise_amd.Context(model, device=0)
        c.set_skin(skin)
        c.set_structure(s, l, x)
        c.md_init(600.0, seed=1)
        c.md_run(10, dt=1.0, T=600.0, tau=50.0)     # warm-up
        c.md_init(600.0, seed=1)
        t0 = time.perf_counter()
        ep, ek = c.md_run(args.md_steps, dt=1.0, T=600.0, tau=50.0)
        dt = time.perf_counter() - t0
        builds = c.skin_builds()
        c.close()
        return ep, ek, dt, builds

    if args.md_skin is not None:  # one variant only (profiling)
        ep, ek, dt, builds = run(args.md_skin)
        ep0, ek0, dt0 = ep, ek, dt
    else:
        ep0, ek0, dt0, _ = run(0.0)
        ep, ek, dt, builds = run(1.0)
    T = ek / (1.5 * len(s) / 11604.0)
    n = len(s)
    return {"workload": f"Cu-Ag 50/50 fcc 12x12x12 ({n} atoms), a=3.8 A, NVT 600 K, dt 1 fs, "
                        f"{args.md_steps} steps; shipped potential nn_CuAg_d3_v1 (145-10-10-1, Rc 7.5 A)",
            "steps_per_s": args.md_steps / dt, "atom_steps_per_s": n * args.md_steps / dt,
            "ms_per_step": 1e3 * dt / args.md_steps,
            "skin_A": 1.0 if args.md_skin is None else args.md_skin, "skin_builds": builds,
            "ms_per_step_lists_every_step": 1e3 * dt0 / args.md_steps,
            "trajectory_equal_to_lists_every_step": bool(np.array_equal(ep0, ep) and np.array_equal(ek0, ek)),
            "T_mean_last_half_K": float(T[len(T) // 2:].mean()),
            "epot_per_atom_first_last_eV": [float(ep[0] / n), float(ep[-1] / n)]}


GA_POT = os.path.join(ROOT, "tests", "golden", "models", "nn_MgCa_d3_v0")


def ga_leg(maise_amd, args):
    """Surrogate-screened evolutionary search, Mg16Ca8 (24 atoms), population 512, crossover + Gaussian mutation
    sigma 0.3 A, seeded, FIRE capped at 300 steps per generation (unconverged structures carried over).
    (1) an audited run: every generation's offspring are also relaxed in full, outside the search clock, for recall
    and error; (2) the same run in compensated BF16 (BF16x3); (3) the same seeded search with and without screening
    for equal search time, at population 512 and at 2048; (4) a sample of offspring relaxed one structure per context
    gives the plain-path rate; (5) the search with the device L-BFGS in place of FIRE (screened, screened in BF16x3,
    every offspring relaxed) and the same sample relaxed one structure per context with it; (6) an audited run that
    ranks offspring after a few FIRE steps of each."""
    from maise_amd.evolve import GAConfig, SurrogateGA, compare_screening, make_offspring
    cfg = GAConfig(pop=args.ga_pop, generations=args.ga_gens, seed=5, relax_cap=300)
    t0 = time.perf_counter()
    ga = SurrogateGA(GA_POT, cfg)
    res = ga.run(log=lambda r: progress(f"audited search generation {r.gen}: best {r.best_eV_atom:.5f} eV/atom")
                 if r.gen % 10 == 0 else None)
    wall = time.perf_counter() - t0
    reps = res.reports[1:]
    ts = float(np.mean([r.seconds_screened for r in reps]))
    tf = float(np.mean([r.seconds_full for r in reps]))
    # plain path: the offspring of one more generation relaxed one structure per context
    pop = [res.best] * 2
    rng = np.random.default_rng(77)
    kids = make_offspring(pop, np.ones(2), cfg, rng, args.ga_single)
    c1 = maise_amd.Context(ga.model, device=0, urep=cfg.urep)
    t1 = time.perf_counter()
    for k in kids:
        c1.set_batch([k])
        c1.relax(max_steps=cfg.max_steps, fmax=cfg.fmax)
    single = len(kids) / (time.perf_counter() - t1)
    t1 = time.perf_counter()
    for k in kids:
        c1.set_batch([k])
        c1.relax_lbfgs(max_steps=cfg.max_steps, fmax=cfg.fmax)
    lb_single = len(kids) / (time.perf_counter() - t1)
    c1.close()
    ga.close()
    # the same screened search with the MLP in compensated BF16 (BF16x3), no audit
    import dataclasses
    cfg_b = dataclasses.replace(cfg, precision=maise_amd.MG_PREC_BF16X3, audit=False)
    gb = SurrogateGA(GA_POT, cfg_b)
    rb = gb.run()
    gb.close()
    tb = float(np.mean([r.seconds_screened for r in rb.reports[1:]]))
    # ranking pre-relaxed offspring: every offspring takes args.ga_prerelax FIRE steps before the surrogate ranks
    # it (audited, half the generations)
    pre = None
    if args.ga_prerelax > 0:
        gp = SurrogateGA(GA_POT, dataclasses.replace(cfg, prerelax=args.ga_prerelax,
                                                     generations=max(1, cfg.generations // 2)))
        rp = gp.run()
        gp.close()
        rr = rp.reports[1:]
        pre = {"prerelax_steps": args.ga_prerelax, "generations": len(rr),
               "recall_top10_mean": float(np.mean([r.recall_top10 for r in rr])),
               "recall_top25_mean": float(np.mean([r.recall_top25 for r in rr])),
               "mae_eV_atom_mean": float(np.mean([r.mae_eV_atom for r in rr])),
               "seconds_per_generation": float(np.mean([r.seconds_screened for r in rr])),
               "seconds_prerelax": float(np.mean([r.seconds_prerelax for r in rr])),
               "seconds_relax_kept": float(np.mean([r.seconds_relax for r in rr]))}
        progress(f"search ranking pre-relaxed offspring: recall@25% {pre['recall_top25_mean']:.2f}")
    # with the device L-BFGS in place of FIRE: screened in FP32, screened in BF16x3, and every offspring relaxed
    lb = {}
    for name, over in (("screened", {}), ("screened_bf16x3", {"precision": maise_amd.MG_PREC_BF16X3}),
                       ("batched_full", {"screen": False})):
        gl = SurrogateGA(GA_POT, dataclasses.replace(cfg, optimizer="lbfgs", audit=False, **over))
        rl = gl.run()
        gl.close()
        lb[name] = cfg.pop / float(np.mean([r.seconds_screened for r in rl.reports[1:]]))
        progress(f"search with L-BFGS, {name}: {lb[name]:.0f} offspring/s")
    def trim(cmp_):
        for name in cmp_["runs"]:
            for r in cmp_["runs"][name]:
                tr = r.pop("trajectory")
                # improvements only (search clock, best eV/atom)
                r["improvements"] = [(round(t, 3), round(e, 5)) for k, (t, e) in enumerate(tr)
                                     if k == 0 or e < tr[k - 1][1] - 1e-9]
        return cmp_

    both = {"unscreened": {"screen": False}, "screened": {"screen": True}}

    def cmp_log(name, seed, r):
        progress(f"comparison run {name} seed {seed}: {r['generations']} generations, best {r['best_eV_atom']:.5f}")
    cmp_ = cmp_large = None
    if args.ga_compare_seconds > 0:
        cmp_ = trim(compare_screening(GA_POT, GAConfig(pop=args.ga_pop, relax_cap=300), args.ga_compare_seconds,
                                      list(range(1, args.ga_compare_seeds + 1)), variants=both, log=cmp_log))
    if args.ga_compare_large_pop > 0:
        # a population whose relaxation batches fill the GPU (relaxation time grows with the batch there)
        cmp_large = trim(compare_screening(GA_POT, GAConfig(pop=args.ga_compare_large_pop, relax_cap=300),
                                           args.ga_compare_large_seconds, list(range(11, 11 + args.ga_compare_seeds + 1)),
                                           variants=both, log=cmp_log))
    n = cfg.pop
    br = {"predict": "seconds_predict", "offspring": "seconds_offspring", "relax_kept": "seconds_relax",
          "train": "seconds_train", "population_update": "seconds_select"}
    return {"workload": f"Mg16Ca8 (24 atoms), shipped potential nn_MgCa_d3_v0 (145-10-10-1, Rc 7.5 A), UREP 100 eV; "
                        f"population {n}, {cfg.generations} generations, seed {cfg.seed}; surrogate 290-64-64-1 on the "
                        "per-species mean descriptor rows, trained online on the GPU; keep top 25% + 10% random; "
                        "FIRE fmax 0.01 eV/A, at most 300 steps per generation (the rest carried over)",
            "offspring_per_s_screened": n / ts, "offspring_per_s_batched_full": n / tf,
            "offspring_per_s_screened_bf16x3": n / tb,
            "offspring_per_s_single_structure": single,
            "screened_over_single": (n / ts) / single, "batched_full_over_single": (n / tf) / single,
            "screened_bf16x3_over_single": (n / tb) / single,
            "relaxations_per_generation_screened": reps[0].relaxed if reps else None,
            "recall_top10_mean": float(np.mean([r.recall_top10 for r in reps])),
            "recall_top25_mean": float(np.mean([r.recall_top25 for r in reps])),
            "recall_top25_last10_mean": float(np.mean([r.recall_top25 for r in reps[-10:]])),
            "recall_random_selection": cfg.keep_top + cfg.keep_random,
            "mae_eV_atom_mean": float(np.mean([r.mae_eV_atom for r in reps])),
            "seconds_per_generation": dict({"screened": ts, "full_relax": tf, "screened_bf16x3": tb},
                                           **{k: float(np.mean([getattr(r, v) for r in reps])) for k, v in br.items()}),
            "train_share_of_generation": float(np.mean([r.seconds_train / r.seconds_screened for r in reps])),
            "carried_per_generation_mean": float(np.mean([r.carried for r in reps])),
            "best_eV_atom_first_last": [res.reports[0].best_eV_atom, res.reports[-1].best_eV_atom],
            "lbfgs": {"offspring_per_s": lb, "offspring_per_s_single_structure": lb_single,
                      "screened_bf16x3_over_plain_fire_single": lb["screened_bf16x3"] / single,
                      "screened_bf16x3_over_plain_lbfgs_single": lb["screened_bf16x3"] / lb_single},
            "screened_prerelaxed": pre,
            "equal_time_comparison": cmp_,
            "equal_time_comparison_large_population": cmp_large,
            "seconds_total": wall}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50,
                    help="untimed evaluations first: the clocks ramp over the first few tens of milliseconds of load")
    ap.add_argument("--n", type=int, default=20, help="fcc supercell edge (4 n^3 atoms)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-leg-cpu", action="store_true", help="skip the CPU figures beside the MD, relax and search legs")
    ap.add_argument("--relax", type=int, default=4096,
                    help="structures in the batched FIRE leg (0 skips it)")
    ap.add_argument("--relax-single", type=int, default=48,
                    help="structures relaxed one at a time for the unbatched comparison")
    ap.add_argument("--md-steps", type=int, default=1000, help="NVT steps of the MD leg (0 skips it)")
    ap.add_argument("--relax-skin", type=float, default=1.0, help="skin (A) of the relax leg's neighbour lists (0 = off)")
    ap.add_argument("--ga-gens", type=int, default=30, help="generations of the screened search (0 skips it)")
    ap.add_argument("--ga-pop", type=int, default=512, help="population of the screened search")
    ap.add_argument("--ga-prerelax", type=int, default=40,
                    help="FIRE steps of every offspring before ranking, in an extra audited run (0 skips it)")
    ap.add_argument("--ga-single", type=int, default=32, help="offspring relaxed one at a time for the plain rate")
    ap.add_argument("--ga-compare-seconds", type=float, default=8.0,
                    help="search clock of each run of the with/without-screening comparison (0 skips it)")
    ap.add_argument("--ga-compare-seeds", type=int, default=3, help="seeds of the with/without-screening comparison")
    ap.add_argument("--ga-compare-large-pop", type=int, default=2048,
                    help="population of a second with/without-screening comparison (0 skips it; one more seed)")
    ap.add_argument("--ga-compare-large-seconds", type=float, default=15.0,
                    help="search clock of each run of the large-population comparison")
    ap.add_argument("--md-skin", type=float, default=None,
                    help="run the MD leg with this skin only (default: 0 and 1.0 A, compared)")
    ap.add_argument("--fire-sweep", type=int, nargs="*", default=None,
                    help="only time FIRE and L-BFGS iterations on batches of this many Mg16Ca8 cells (e.g. 64 179 512 "
                         "2048), and both to convergence on up to 179 cells and offspring")
    ap.add_argument("--fire-iters", type=int, default=300, help="iterations per batch of --fire-sweep")
    ap.add_argument("--rates-only", action="store_true",
                    help="only run the library's issue-rate microbenchmark (FP32 FMA, exp2, cos, sqrt, rcp) and print it")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 256 MiB (5 times): the reference point that scales the "
                         "FETCH_SIZE/WRITE_SIZE counters of a profiling run to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); the multi-GPU sharding is "
              "design-only.", file=sys.stderr)
        sys.exit(2)
    batch = None
    if args.relax > 0:
        from maise_amd import workloads as wl0
        # built before the GPU is touched: the generator forks worker processes
        batch = wl0.cupdag_batch(args.relax, seed=3)
    if args.copy_calibration:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4, "dispatches": 5}))
        return

    import maise_amd
    from maise_amd import workloads as wl
    if args.fire_sweep:
        # FIRE iterations of batches of Mg16Ca8 cells (24 atoms): wall time per iteration against batch size, every
        # structure active throughout (fmax below reach)
        from maise_amd.evolve import GAConfig, random_structure
        cfg = GAConfig()
        rng = np.random.default_rng(3)
        cells = [random_structure(cfg, rng) for _ in range(max(args.fire_sweep))]
        c = maise_amd.Context(maise_amd.Model(GA_POT), device=0, urep=100.0)
        c.set_batch(cells[:16])
        c.relax(max_steps=20, fmax=0.01)
        out = {}
        c.relax_lbfgs(max_steps=20, fmax=0.01)
        out, out_lb = {}, {}
        for n in args.fire_sweep:
            for opt, dst in ((c.relax, out), (c.relax_lbfgs, out_lb)):
                c.set_batch(cells[:n])
                t0 = time.perf_counter()
                opt(max_steps=args.fire_iters, fmax=1e-9)
                dst[n] = 1e6 * (time.perf_counter() - t0) / args.fire_iters
        # to convergence (fmax 0.01 eV/A), FIRE against L-BFGS: random cells, and crossover offspring of 8 relaxed
        # parents (the search's operators, seed 5)
        from maise_amd.evolve import make_offspring
        nconv = min(179, max(args.fire_sweep))
        c.set_batch(cells[:nconv])
        c.relax(max_steps=3000, fmax=0.01)
        par = [(cells[k][0], cells[k][1], c.result(k, forces=False)["X"]) for k in range(min(8, nconv))]
        kids = make_offspring(par, np.ones(len(par)), cfg, np.random.default_rng(5), nconv)
        conv = {}
        for name, batch in (("random", cells[:nconv]), ("offspring", kids)):
            for opt in ("fire", "lbfgs"):
                c.set_batch(batch)
                t0 = time.perf_counter()
                st, dn = (c.relax if opt == "fire" else c.relax_lbfgs)(max_steps=3000, fmax=0.01)
                conv[f"{name}_{opt}"] = {"seconds": time.perf_counter() - t0, "steps_mean": float(np.mean(st)),
                                         "steps_max": int(np.max(st)), "converged": int(np.sum(dn))}
        c.close()
        print(json.dumps({"workload": f"FIRE and L-BFGS (m = 20) on batches of random Mg16Ca8 cells, nn_MgCa_d3_v0, "
                                      f"UREP 100 eV, {args.fire_iters} iterations",
                          "us_per_iteration_by_cells": out, "lbfgs_us_per_iteration_by_cells": out_lb,
                          "to_convergence_cells": nconv, "to_convergence": conv}))
        return
    if args.rates_only:
        r = measured_rates(maise_amd)
        print(json.dumps({"ops_per_s": r, "fma_TFLOPs": 2 * r["fma"] / 1e12,
                          "exp2_over_fma": r["exp2"] / r["fma"], "cos_over_fma": r["cos"] / r["fma"]}))
        return
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is optional plumbing here
        sync = lambda: None  # noqa: E731

    spc, lat, x = wl.cu_fcc(args.n, 3.61, 0.05, seed=1)
    N = len(spc)
    model = maise_amd.Model(POT)
    ctx = maise_amd.Context(model, device=0)
    ctx.set_structure(spc, lat, x)
    for _ in range(args.warmup):
        ctx.eval(forces=True)
    ctx.set_timing(True)  # per-stage HIP events on the context's stream; evaluations stay back to back
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.eval(forces=True)
    # waits for the last evaluation and checks its flags (a capacity overflow would be evaluated again here, inside
    # the timed region); stage means over the K timed evaluations
    stages = ctx.timing().astype(np.float64)
    sync()
    dt = time.perf_counter() - t0
    res = ctx.result(0)
    M, K = algorithmic_counts(ctx, lat, model.rc)  # from the neighbour list of the timed evaluations
    ms = 1e3 * dt / args.steps
    value = N * args.steps / dt

    fl_desc, fl_force = flops_model(M, K)
    t_force = stages[3] * 1e-3
    t_desc = stages[1] * 1e-3
    # FP32 with shared, non-negative etas: the force kernel is the split-FP16 matrix-core form
    fname = "k_force_mh" if model.shared_eta else "k_force"
    kernel, fl, t = (fname, fl_force, t_force) if t_force >= t_desc else ("k_desc", fl_desc, t_desc)
    achieved = fl / t / 1e12
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_cu32k_w64_fp32.json")
    if os.path.exists(tpath) and N == 32000:
        tj = json.load(open(tpath))
        kk = tj["kernels_KiB_per_dispatch"].get(kernel)
        if kk:
            cal = tj["calibration"]
            traffic = int(1024 * (kk["FETCH_SIZE"] * cal["fetch_scale"] + kk["WRITE_SIZE"] * cal["write_scale"]))
            traffic_src = "profiles/traffic_cu32k_w64_fp32.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated)"
    rates = measured_rates(maise_amd)  # after the timed region: the microbenchmark runs on its own
    model_roof = roofline_model(N, M, K, stages, ms, model, rates)
    roof = {"bound": "valu",
            "bound_detail": "FP32 vector ALU issue, FMA and transcendental (neither HBM nor the matrix cores bound "
                            "the descriptor and force kernels; the force kernel's triplet contraction runs on the "
                            "matrix cores beside it); priced at the published FP32 peak, which counts packed FMA "
                            "(packed FP32 is not used; the measured non-packed FMA issue limit is in "
                            "measured_rates_per_s)",
            "kernel": kernel, "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic, "traffic_source": traffic_src,
            "frac_of_measured_fma_issue": round(achieved / (2 * rates["fma"] / 1e12), 4),
            "flops_per_launch": fl, "kernel_ms": round(t * 1e3, 4), "sum_M": M, "sum_K": K}
    roof.update(model_roof)

    # FP32 path against the reference's own FP64 output (CALL_MAISE compiled from its sources, committed as a golden
    # vector) on the 256-atom cell with the same potential
    err = None
    try:
        z = np.load(os.path.join(ROOT, "tests", "golden", "ref_cases.npz"))
        k = "config1_w64"
        c1 = maise_amd.Context(maise_amd.Model(os.path.join(ROOT, str(z[k + "/model"]))))
        c1.set_structure(z[k + "/spc"], z[k + "/lat"], z[k + "/x"], nd=int(z[k + "/nd"]), xt=int(z[k + "/xt"]))
        g = c1.evaluate()
        n1 = len(z[k + "/spc"])
        err = {"case": "Cu fcc 4x4x4 (256 atoms), same potential, reference FP64 output",
               "dE_per_atom_eV": abs(g["E"] - float(z[k + "/E"])) / n1,
               "dF_max_eV_A": float(np.abs(g["F"] - z[k + "/F"].reshape(-1, 3)).max()),
               "dU_max_eV_per_atom": float(np.abs(g["U"] - z[k + "/U"]).max() / n1)}
        c1.close()
    except Exception as e:  # informative only
        err = {"unavailable": str(e)[:200]}

    # same workload with the BF16 MLP forms on MFMA: compensated (BF16x3, every product) and one pass in the
    # hidden x hidden product
    def bf16_line(prec):
        try:
            cb = maise_amd.Context(model, device=0, precision=prec)
            cb.set_structure(spc, lat, x)
            for _ in range(args.warmup):
                cb.eval(forces=True)
            sync()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                cb.eval(forces=True)
            sync()
            dtb = time.perf_counter() - t0
            rb = cb.result(0)
            cb.close()
            return {"value": N * args.steps / dtb, "ms_per_step": 1e3 * dtb / args.steps,
                    "dE_per_atom_vs_fp32_eV": abs(rb["E"] - res["E"]) / N,
                    "dF_max_vs_fp32_eV_A": float(np.abs(rb["F"] - res["F"]).max())}
        except Exception as e:
            return {"unavailable": str(e)[:200]}

    bf = bf16_line(maise_amd.MG_PREC_BF16X3)
    bf1 = bf16_line(maise_amd.MG_PREC_BF16)

    progress(f"headline done: {value:.4g} atom E+F evals/s, {ms:.4f} ms per evaluation")
    relax = None
    if batch is not None:
        try:
            relax = relax_leg(maise_amd, batch, args)
            progress(f"relax leg done: {relax['relaxed_structures_per_s']:.1f} relaxed structures/s")
        except Exception as e:
            relax = {"unavailable": str(e)[:300]}
    md = None
    if args.md_steps > 0:
        try:
            md = md_leg(maise_amd, args)
            progress(f"md leg done: {md['ms_per_step']:.4f} ms per step")
        except Exception as e:
            md = {"unavailable": str(e)[:300]}
    ga = None
    if args.ga_gens > 0:
        try:
            ga = ga_leg(maise_amd, args)
            progress("search leg done")
        except Exception as e:
            ga = {"unavailable": str(e)[:300]}

    cpu = None
    if not args.no_cpu:
        # the host cores this process may use, at most the 16-core share of one GPU on a shared node
        threads = min(host_cpu_info()[0], 16)
        try:
            cpu = cpu_baseline(threads)
            progress(f"cpu baseline done: {cpu['value']:.1f} atom E+F evals/s on {threads} threads")
        except Exception as e:
            cpu = {"unavailable": str(e)[:200]}
        if not args.no_leg_cpu:
            try:
                legs = leg_cpu_baselines(threads, batch if relax is not None else None)
                for key, leg in (("md", md), ("relax", relax), ("search", ga)):
                    if key in legs and isinstance(leg, dict):
                        leg["cpu_baseline"] = legs[key]
                progress("leg CPU baselines done")
            except Exception as e:
                progress(f"leg CPU baselines unavailable: {str(e)[:200]}")

    out = {
        "metric": "atom E+F evals/s", "value": value, "unit": "atom E+F evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"Cu fcc {args.n}x{args.n}x{args.n} ({N} atoms), a=3.61 A, jitter 0.05 A, "
                               "synthetic 51-64-64-1 tanh potential, Rc 6.0 A, FP32 E+F+virial",
                   "atoms": N, "parallelism": "single GPU"},
        "stages_ms": {k: round(float(v), 4) for k, v in zip(("neighbours", "descriptors", "mlp", "forces",
                                                             "assemble", "total"), stages)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu and "value" in cpu else None,
        "error_fp64": err,
        "bf16x3_mlp": bf,
        "bf16_hidden_mlp": bf1,
        "relax": relax,
        "md": md,
        "screened_search": ga,
        "energy_eV": res["E"],
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
