"""Generate the golden fixtures and synthetic potentials (run in the container that holds the reference).

Outputs
  maise_amd/potentials/*.nn      seeded synthetic models in the nn_* format (input ranges calibrated with the
                                 FP64 oracle on the workload structure, so inputs are O(1))
  tests/golden/ref_cases.npz     golden vectors from the reference's own E/F/S path compiled from its sources
                                 (oracle/_ref/maise_ref, NP=1): inputs, H, E, EA, F, U, neighbour sets, G, X
  tests/golden/parse_descriptors.npz   inputs and descriptor values of the reference regression fixture
                                 test/parse (POSCAR.0 files + ref/e* descriptor files, 82-wide Cu-Ag basis)
  tests/golden/models/           copies of the reference model files the fixtures use (data)

usage: python tests/golden/make_golden.py [--skip-ref]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from maise_amd import modelfile, poscar, workloads as wl  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

REF = "/root/reference"
MODELS = os.path.join(HERE, "models")
POT = os.path.join(ROOT, "maise_amd", "potentials")


def oracle_descriptors(path, spc, lat, x, nd=3):
    m = po.OracleModel(path)
    return m.evaluate(spc, lat, x, nd=nd, forces=False, want_g=True)["G"]


def make_potentials():
    os.makedirs(POT, exist_ok=True)
    out = {}
    # config 1 / 4 structures: Cu fcc a=3.61, sigma=0.05 jitter (seed 1)
    c1 = wl.cu_fcc(4, 3.61, 0.05, seed=1)
    for name, hidden, seed in (("cu_51_10_10.nn", (10, 10), 7), ("cu_51_64_64.nn", (64, 64), 8)):
        p = os.path.join(POT, name)
        wl.calibrated_model(p, [29], hidden, seed, [c1], oracle_descriptors)
        out[name] = p
    # config 2: Cu-Ag 50/50 fcc a=3.8 (species-resolved 145 inputs)
    c2 = wl.alloy_fcc(4, 3.8, 2, 0.05, seed=2)
    p = os.path.join(POT, "cuag_145_10_10.nn")
    wl.calibrated_model(p, [29, 47], (10, 10), 9, [c2], oracle_descriptors)
    out["cuag_145_10_10.nn"] = p
    # config 3: Cu-Pd-Ag ternary (282 inputs), calibrated on a few random triclinic cells
    rng = np.random.default_rng(11)
    c3 = [wl.random_triclinic(int(rng.integers(8, 65)), 3, rng) for _ in range(6)]
    p = os.path.join(POT, "cupdag_282_10_10.nn")
    wl.calibrated_model(p, [29, 46, 47], (10, 10), 10, c3, oracle_descriptors)
    out["cupdag_282_10_10.nn"] = p
    # species-specific bases (test data, kept with the fixtures)
    cs = [wl.alloy_fcc(3, 3.8, 2, 0.1, seed=7)]
    for name, mk in (("species_generic", modelfile.generic_basis), ("species_std", modelfile.standard_basis)):
        p = os.path.join(MODELS, f"nn_CuAg_{name}")
        wl.calibrated_model(p, [29, 47], (10, 10), 11, cs, oracle_descriptors,
                            basis=[mk(rc=6.0), mk(rc=6.0, scale=1.1)])
        out[name] = p
    return out


def ref_case(name, model, spc, lat, x, nd=3, xt=1, urep=0.0, p_gpa=0.0, keep_g=True):
    r = po.run_reference(model, spc, lat, x, nd=nd, xt=xt, nm=1000, urep=urep, p_gpa=p_gpa)
    rel = os.path.relpath(model, ROOT)
    d = {f"{name}/model": np.array(rel), f"{name}/spc": np.asarray(spc, np.int32),
         f"{name}/lat": np.asarray(lat, np.float64).reshape(3, 3), f"{name}/x": np.asarray(x, np.float64),
         f"{name}/nd": np.array(nd), f"{name}/xt": np.array(xt), f"{name}/urep": np.array(urep),
         f"{name}/p_gpa": np.array(p_gpa), f"{name}/H": np.array(r["H"]), f"{name}/E": np.array(r["E"]),
         f"{name}/EA": r["EA"], f"{name}/F": r["F"], f"{name}/U": r["U"], f"{name}/nn": r["nn"].astype(np.int32),
         f"{name}/nbr": r["nbr"].astype(np.int32), f"{name}/X": r["X"]}
    if keep_g:
        d[f"{name}/G"] = r["G"]
    print(f"  {name}: N={len(spc)} H={r['H']:.10f} M={int(r['nn'].sum())}", flush=True)
    return d


def make_ref_cases(pots):
    cases = {}
    g = lambda n: os.path.join(MODELS, n)
    # interface cluster (interface/c/c2maise.c:42-51)
    lat = np.eye(3) * 20.0
    x = np.array([[2, 0, 0], [-2, 0, 0], [0, 2, 0], [0, -2, 0]], float)
    cases.update(ref_case("iface", g("nn_CuAg_d3_v0_iface"), np.zeros(4, np.int32), lat, x, nd=0))
    # relax fixture start: L1_2 Cu3Pd, a = 4.0, fractional input (test/relax/POSCAR)
    pc = poscar.read_poscar(os.path.join(REF, "test", "relax", "POSCAR"))
    cases.update(ref_case("relax", g("nn_CuPd_d3_vp_relax"), pc.species, pc.lattice, pc.frac, nd=3, xt=0))
    cases.update(ref_case("relax_p", g("nn_CuPd_d3_vp_relax"), pc.species, pc.lattice, pc.frac, nd=3, xt=0, p_gpa=5.0))
    # jittered elemental and alloys with the shipped models (Rc 7.5 A)
    spc, lat, x = wl.cu_fcc(3, 3.61, 0.1, seed=3)
    cases.update(ref_case("cu", g("nn_Cu_d3_v1"), spc, lat, x))
    spc, lat, x = wl.alloy_fcc(3, 3.8, 2, 0.1, seed=3)
    cases.update(ref_case("cuag", g("nn_CuAg_d3_v1"), spc, lat, x))
    rng = np.random.default_rng(5)
    spc, lat, x = wl.random_triclinic(20, 3, rng)
    cases.update(ref_case("cupdag_tri", g("nn_CuPdAg_d3_v1"), spc, lat, x))
    spc, lat, x = wl.random_triclinic(12, 2, rng, dmin=1.6)
    cases.update(ref_case("urep", g("nn_CuAg_d3_v1"), spc, lat, x, urep=100.0))
    spc, lat, x = wl.random_triclinic(8, 3, rng)
    cases.update(ref_case("small_tri", g("nn_CuPdAg_d3_v1"), spc, lat, x))
    spc, lat, x = wl.random_triclinic(24, 2, np.random.default_rng(21))
    cases.update(ref_case("mgca", g("nn_MgCa_d3_v0"), spc, lat, x))
    # synthetic models at the workload cutoff 6.0 A
    spc, lat, x = wl.cu_fcc(4, 3.61, 0.05, seed=1)
    cases.update(ref_case("config1", pots["cu_51_10_10.nn"], spc, lat, x))
    cases.update(ref_case("config1_w64", pots["cu_51_64_64.nn"], spc, lat, x, keep_g=False))
    # Cu-Ag at a=3.8: the 5th shell (6.0083 A) straddles Rc=6.0 under small jitter -> borderline sets
    spc, lat, x = wl.alloy_fcc(3, 3.8, 2, 0.004, seed=4)
    cases.update(ref_case("shell", pots["cuag_145_10_10.nn"], spc, lat, x))
    rng = np.random.default_rng(12)
    spc, lat, x = wl.random_triclinic(40, 3, rng)
    cases.update(ref_case("config3_like", pots["cupdag_282_10_10.nn"], spc, lat, x))
    # species-specific basis blocks (Ag block scaled by 1.1): cross-species etas differ by pair (the mixing rule
    # of src/nprs.c:233-252), once with a zeta = {1, 3} basis matching neither shipped G4 table, once with the
    # 43-function table
    spc, lat, x = wl.alloy_fcc(3, 3.8, 2, 0.1, seed=8)
    cases.update(ref_case("species_generic", pots["species_generic"], spc, lat, x))
    cases.update(ref_case("species_std", pots["species_std"], spc, lat, x))
    np.savez_compressed(os.path.join(HERE, "ref_cases.npz"), **cases)


def make_parse_fixture():
    """test/parse: POSCAR.0 inputs and the reference's descriptor dumps ref/e* (EFS=1, 82-wide Cu-Ag basis)."""
    base = os.path.join(REF, "test", "parse")
    idx = {}
    with open(os.path.join(base, "ref", "index.dat")) as f:
        for line in f:
            t = line.split()
            if len(t) >= 2:
                idx[t[0]] = t[1]
    basis = modelfile.read_basis(os.path.join(base, "basis"))
    net = modelfile.random_network([29, 47], [82, 10, 10, 1], seed=1)
    mpath = os.path.join(MODELS, "nn_CuAg_parse_basis")
    modelfile.write_model(mpath, net, basis)
    out = {}
    n = 0
    for ename in sorted(idx):
        epath = os.path.join(base, "ref", ename)
        if not os.path.exists(epath):
            continue
        with open(epath) as f:
            lines = f.read().split("\n")
        hdr = lines[0].split()
        efs, D, N = int(hdr[0]), int(hdr[1]), int(hdr[2])
        li = 1
        if efs in (1, 3):
            li += 1
        if efs in (2, 3):
            li += 1
        li += 1  # path
        spc = np.zeros(N, np.int32)
        G = np.zeros((N, D))
        for i in range(N):
            spc[i] = int(lines[li]); li += 1
            for k in range(D):
                G[i, k] = float(lines[li]); li += 1
        pc = poscar.read_poscar(os.path.join(base, idx[ename], "POSCAR.0"))
        if not np.array_equal(pc.species, spc):
            raise RuntimeError(f"{ename}: species mismatch")
        out[f"{ename}/lat"] = pc.lattice
        out[f"{ename}/frac"] = pc.frac
        out[f"{ename}/spc"] = spc
        out[f"{ename}/G"] = G
        n += 1
    np.savez_compressed(os.path.join(HERE, "parse_descriptors.npz"), **out)
    print(f"  parse fixture: {n} structures")


if __name__ == "__main__":
    os.makedirs(MODELS, exist_ok=True)
    pots = make_potentials()
    print("potentials:", sorted(pots))
    make_parse_fixture()
    if "--skip-ref" not in sys.argv:
        make_ref_cases(pots)
