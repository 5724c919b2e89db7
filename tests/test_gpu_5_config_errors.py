"""FP32 path against the FP64 oracle on the states the benchmark workloads produce.

  * MD: the Cu-Ag 50/50 fcc alloy (a = 3.8 A) under NVT at 600 K with the shipped nn_CuAg_d3_v1 (145-10-10-1,
    Rc 7.5 A), Maxwell-Boltzmann start seed 1, dt 1 fs: at each snapshot the positions the device holds are
    evaluated by the oracle and compared with the device's own energy, forces and virial.
  * Batched relaxation: Cu-Pd-Ag structures of the batch generator (seed 3, 8-64 atoms, random triclinic cells)
    relaxed by FIRE to 0.01 eV/A with nn_CuPdAg_d3_v1 and UREP 100 eV, every structure in one launch; the relaxed
    positions are evaluated by the oracle (energy, forces, and the oracle's own largest force at the end point).
  * Search: random Mg16Ca8 cells of the evolutionary search's initial population (24-32 A^3/atom, dmin 2 A) relaxed
    the same way with nn_MgCa_d3_v0 and UREP 100 eV.

Bars (FP32 path, FP64 reference): 1e-5 eV/atom, 1e-4 eV/A, 1e-4 eV/atom virial.  Default sizes keep the oracle
to seconds: MD 6x6x6 (864 atoms) to step 100, 64 relaxed Cu-Pd-Ag structures, 32 Mg16Ca8 cells.
MAISE_FULL_REPORT=1 runs the MD workload itself (12x12x12, 6912 atoms, snapshots every 250 steps to 1000) and
MAISE_REPORT_FILE=<path> appends the tables there.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

E_TOL, F_TOL, U_TOL = 1e-5, 1e-4, 1e-4


def _report(title, header, rows):
    lines = [f"\n{title}", "| " + " | ".join(header) + " |", "|" + "---|" * len(header)]
    lines += ["| " + " | ".join(r) + " |" for r in rows]
    text = "\n".join(lines)
    print(text, flush=True)
    path = os.environ.get("MAISE_REPORT_FILE")
    if path:
        with open(path, "a") as f:
            f.write(text + "\n")


def test_md_alloy_fp32_vs_fp64_along_nvt():
    import maise_amd
    from maise_amd import workloads as wl
    from oracle import pyoracle as po
    full = os.environ.get("MAISE_FULL_REPORT") == "1"
    n, snaps = (12, [0, 250, 500, 750, 1000]) if full else (6, [0, 50, 100])
    path = os.path.join(GOLDEN, "models", "nn_CuAg_d3_v1")
    s, l, x = wl.alloy_fcc(n, 3.8, 2, 0.0, seed=2)
    N = len(s)
    print(f"[gpu case] MD alloy {N} atoms, snapshots {snaps}", flush=True)
    c = maise_amd.Context(maise_amd.Model(path))
    c.set_structure(s, l, x)
    c.md_init(600.0, seed=1)
    om = po.OracleModel(path)
    rows, done = [], 0
    for step in snaps:
        if step > done:
            c.md_run(step - done, dt=1.0, T=600.0, tau=50.0)
            done = step
        else:
            c.eval(forces=True)
        r = c.result(0)
        o = om.evaluate(s, l, r["X"], nd=3, xt=1, threshold=False)
        dE = abs(r["E"] - o["E"]) / N
        dF = np.abs(r["F"] - o["F"])
        dU = np.abs(r["U"] - o["U"]).max() / N
        rows.append([str(step), f"{dE:.2e}", f"{dF.max():.2e}", f"{np.sqrt(np.mean(dF ** 2)):.2e}", f"{dU:.2e}"])
        assert dE < E_TOL and dF.max() < F_TOL and dU < U_TOL, rows[-1]
    c.close()
    _report(f"Cu-Ag 50/50 fcc {n}x{n}x{n} ({N} atoms), NVT 600 K, nn_CuAg_d3_v1: FP32 path vs FP64 oracle",
            ["step", "dE/atom (eV)", "max |dF| (eV/A)", "RMS dF (eV/A)", "max |dU|/atom (eV)"], rows)


def _relaxed_vs_oracle(model_path, structs, urep, title):
    import maise_amd
    from oracle import pyoracle as po
    c = maise_amd.Context(maise_amd.Model(model_path), urep=urep)
    c.set_batch(structs)
    steps, conv = c.relax(max_steps=3000, fmax=0.01)
    assert conv.all(), steps
    om = po.OracleModel(model_path)
    worst = np.zeros(4)
    for k, (spc, lat, _) in enumerate(structs):
        r = c.result(k)
        n = len(spc)
        o = om.evaluate(spc, lat, r["X"], nd=3, xt=1, urep=urep, threshold=False)
        fo = np.linalg.norm(o["F"], axis=1).max()
        errs = [abs(r["E"] - o["E"]) / n, np.abs(r["F"] - o["F"]).max(), np.abs(r["U"] - o["U"]).max() / n, fo]
        worst = np.maximum(worst, errs)
    c.close()
    _report(title, ["structures", "FIRE steps mean / max", "max dE/atom (eV)", "max |dF| (eV/A)",
                    "max |dU|/atom (eV)", "oracle max |F| at the end points (eV/A)"],
            [[str(len(structs)), f"{steps.mean():.0f} / {steps.max()}", f"{worst[0]:.2e}", f"{worst[1]:.2e}",
              f"{worst[2]:.2e}", f"{worst[3]:.4f}"]])
    assert worst[0] < E_TOL and worst[1] < F_TOL and worst[2] < U_TOL
    assert worst[3] < 0.01 + F_TOL  # the oracle agrees that every end point meets the force threshold


def test_batched_relax_fp32_vs_fp64():
    from maise_amd import workloads as wl
    print("[gpu case] batched relax, 64 Cu-Pd-Ag structures", flush=True)
    structs = wl.cupdag_batch(64, seed=3)
    _relaxed_vs_oracle(os.path.join(GOLDEN, "models", "nn_CuPdAg_d3_v1"), structs, 100.0,
                       "Cu-Pd-Ag batch (seed 3, first 64 structures), FIRE to 0.01 eV/A, nn_CuPdAg_d3_v1, UREP 100 eV: "
                       "relaxed states, FP32 path vs FP64 oracle")


def test_search_cells_fp32_vs_fp64():
    from maise_amd.evolve import GAConfig, random_structure
    print("[gpu case] Mg16Ca8 cells, 32 structures", flush=True)
    cfg = GAConfig()
    rng = np.random.default_rng(cfg.seed)
    structs = [random_structure(cfg, rng) for _ in range(32)]
    _relaxed_vs_oracle(os.path.join(GOLDEN, "models", "nn_MgCa_d3_v0"), structs, cfg.urep,
                       "Mg16Ca8 random cells (search's initial population, seed 5, 32 structures), FIRE to 0.01 eV/A, "
                       "nn_MgCa_d3_v0, UREP 100 eV: relaxed states, FP32 path vs FP64 oracle")
