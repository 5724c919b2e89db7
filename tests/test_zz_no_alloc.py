"""GPU: mg_eval_static, the evaluation that allocates nothing (all intermediate buffers in the library's fixed
per-device scratch array), against the existing entry point (mg_set_structure + mg_eval + mg_get) on the same
inputs, to the FP32 path's bars: energy 1e-5 eV/atom, forces 1e-4 eV/A, virial 1e-4 eV/atom.  Two calls in a row,
the second on a smaller structure (the scratch keeps the first call's contents), then a second model.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
E_TOL, F_TOL, U_TOL = 1e-5, 1e-4, 1e-4


def _compare(model, spc, lat, x, label):
    import maise_amd
    print(f"[gpu case] no-alloc {label}: N={len(spc)}", flush=True)
    s = maise_amd.eval_static(model, spc, lat, x)
    c = maise_amd.Context(model)
    c.set_structure(spc, lat, x)
    r = c.evaluate()
    EA = c.result()["EA"]
    c.close()
    N = len(spc)
    de = abs(s["E"] - r["E"]) / N
    dea = np.abs(s["EA"] - EA).max()
    df = np.abs(s["F"] - r["F"]).max()
    du = np.abs(s["U"] - r["U"]).max() / N
    print(f"{label}: dE/atom={de:.2e} dEA={dea:.2e} dF={df:.2e} dU/atom={du:.2e}", flush=True)
    assert np.isfinite(s["F"]).all()
    assert de < E_TOL and dea < 10 * E_TOL and df < F_TOL and du < U_TOL


def test_static_eval_matches_context_two_calls():
    import maise_amd
    from maise_amd import workloads as wl
    m = maise_amd.Model(os.path.join(ROOT, "maise_amd", "potentials", "cu_51_64_64.nn"))
    spc, lat, x = wl.cu_fcc(6, 3.61, 0.05, seed=21)          # 864 atoms
    _compare(m, spc, lat, x, "cu 864 (51-64-64-1)")
    spc, lat, x = wl.cu_fcc(4, 3.61, 0.05, seed=22)          # then 256 atoms: smaller than the first call
    _compare(m, spc, lat, x, "cu 256 (51-64-64-1)")


def test_static_eval_second_model_ternary():
    import maise_amd
    from maise_amd import workloads as wl
    m = maise_amd.Model(os.path.join(ROOT, "maise_amd", "potentials", "cupdag_282_10_10.nn"))
    rng = np.random.default_rng(5)
    spc, lat, x = wl.random_triclinic(40, 3, rng)
    _compare(m, spc, lat, x, "cupdag 40 triclinic (282-10-10-1)")
