"""Batched L-BFGS relaxation on the device (mg_relax_lbfgs).

The reference relaxes with GSL's vector_bfgs2 (CELL_MIN, src/cmin.c:203-245), which is absent here, so trajectory
parity with the reference is not a bar.  Checked instead: the device recursion against the oracle's FP64
restatement of the same method (oracle/pyoracle.py: lbfgs_relax, itself pinned by the explicit BFGS update in
tests/test_lbfgs_oracle.py) over the first steps, where FP32 forces keep the two within a small distance; every
structure ending below the force threshold with the FP64 oracle agreeing at the relaxed positions; results that do
not depend on graph replay, on buffers outgrowing a capture (checkpoint restore of the history), on skin lists, on
the batch being re-packed as structures converge, or on the structure being relaxed alone or in a batch.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

POT = os.path.join(GOLDEN, "models", "nn_CuPdAg_d3_v1")


def _min_image(d, lat):
    f = d @ np.linalg.inv(lat)
    return (f - np.rint(f)) @ lat


def test_lbfgs_first_steps_follow_the_oracle():
    """Six L-BFGS steps on the device and in the FP64 restatement from the same start: positions agree to 2e-4 A
    (the device forces are FP32); a wrong sign, loop order or pair slot in the recursion moves them by 1e-2 A or
    more."""
    import maise_amd
    from maise_amd import workloads as wl
    from oracle import pyoracle as po
    rng = np.random.default_rng(31)
    structs = [wl.random_triclinic(int(rng.integers(8, 33)), 3, rng) for _ in range(4)]
    print("[gpu case] L-BFGS first 6 steps vs oracle, 4 cells", flush=True)
    m = maise_amd.Model(POT)
    om = po.OracleModel(POT)
    for nsteps in (1, 6):
        c = maise_amd.Context(m, urep=100.0)
        c.set_batch(structs)
        steps, done = c.relax_lbfgs(max_steps=nsteps, fmax=1e-6)
        assert (steps == nsteps).all() and not done.any()
        for k, (spc, lat, x) in enumerate(structs):
            X = c.result(k, forces=False)["X"]
            xo = po.lbfgs_relax(om, spc, lat, x, fmax=1e-6, max_steps=50, urep=100.0, return_x_after=nsteps)
            d = _min_image(X - xo, np.asarray(lat).reshape(3, 3))
            assert np.abs(d).max() < 2e-4, (nsteps, k, np.abs(d).max())
        c.close()


def test_lbfgs_batch_relaxes_and_matches_oracle():
    """Every structure converges; the FP64 oracle confirms |F| < fmax and the energy at the relaxed positions; the
    restatement on FP64 forces reaches the same energy per atom; L-BFGS takes fewer steps than FIRE on the batch;
    a structure relaxed alone takes the same steps to the same positions."""
    import maise_amd
    from maise_amd import workloads as wl
    from oracle import pyoracle as po
    rng = np.random.default_rng(31)
    structs = [wl.random_triclinic(int(rng.integers(8, 33)), 3, rng) for _ in range(12)]
    print("[gpu case] L-BFGS batch of 12", flush=True)
    m = maise_amd.Model(POT)
    c = maise_amd.Context(m, urep=100.0)
    c.set_batch(structs)
    steps, done = c.relax_lbfgs(max_steps=2000, fmax=0.01)
    assert done.all(), steps
    res = c.batch_result()
    X = [c.result(k)["X"] for k in range(len(structs))]
    om = po.OracleModel(POT)
    for k, (spc, lat, _) in enumerate(structs):
        n = len(spc)
        o = om.evaluate(spc, lat, X[k], urep=100.0)
        assert np.abs(o["F"]).max() < 0.011
        assert res["E"][k] == pytest.approx(o["E"], abs=1e-5 * n)
    for k in (0, 5):
        spc, lat, x = structs[k]
        _, e_cpu, _, ok = po.lbfgs_relax(om, spc, lat, x, fmax=0.01, urep=100.0)
        assert ok
        assert res["E"][k] / len(spc) == pytest.approx(e_cpu / len(spc), abs=2e-3)
    cf = maise_amd.Context(m, urep=100.0)
    cf.set_batch(structs)
    sf, df = cf.relax(max_steps=3000, fmax=0.01)
    assert df.all()
    print(f"  mean steps: L-BFGS {steps.mean():.1f}, FIRE {sf.mean():.1f}", flush=True)
    assert steps.mean() < 0.8 * sf.mean()
    cf.close()
    c1 = maise_amd.Context(m, urep=100.0)
    c1.set_batch([structs[3]])
    s1, d1 = c1.relax_lbfgs(max_steps=2000, fmax=0.01)
    assert d1[0] and s1[0] == steps[3]
    assert c1.batch_result()["E"][0] == pytest.approx(res["E"][3], abs=1e-9)
    assert np.abs(c1.result(0)["X"] - X[3]).max() < 1e-9
    c.close()
    c1.close()


def test_lbfgs_graph_replay_matches_synchronised_loop():
    """Graph with default headroom, graph with zero headroom (rows outgrow the capture, the checkpoint including the
    L-BFGS history is restored and the iterations redone) and every iteration synchronised: identical results."""
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] L-BFGS graph vs synchronised, batch of 10", flush=True)
    rng = np.random.default_rng(44)
    structs = [wl.random_triclinic(int(rng.integers(8, 40)), 3, rng) for _ in range(10)]
    m = maise_amd.Model(POT)
    out = []
    for enable, headroom in ((True, 0.125), (True, 0.0), (False, 0.125)):
        c = maise_amd.Context(m, urep=100.0)
        c.set_graph(enable, headroom)
        c.set_batch(structs)
        steps, done = c.relax_lbfgs(max_steps=1500, fmax=0.01)
        res = c.batch_result()
        X = np.concatenate([c.result(k, forces=False)["X"] for k in range(len(structs))])
        out.append((steps, done, res["E"], res["F"], X))
        c.close()
    for k in (1, 2):
        for a, b in zip(out[0], out[k]):
            assert np.array_equal(a, b)
    assert out[0][1].all()


def test_lbfgs_repacked_batch_with_skin_lists_matches_single_structures():
    """48 structures converge at different steps, so the batch is re-packed several times (history rings, previous
    forces and 1/(y.s) gathered into the new layout on the device); with skin lists on and off the results are
    bit-identical, and structures relaxed one per context take the same steps to the same positions."""
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] L-BFGS 48 Cu-Pd-Ag, re-packing, skin on/off, singles", flush=True)
    structs = wl.cupdag_batch(48, seed=3, workers=1)
    m = maise_amd.Model(POT)
    runs = []
    for skin in (0.0, 1.0):
        c = maise_amd.Context(m, urep=100.0)
        c.set_skin(skin)
        c.set_batch(structs)
        steps, done = c.relax_lbfgs(max_steps=3000, fmax=0.01)
        log = c.relax_log()
        res = c.batch_result()
        X = [c.result(k, forces=False)["X"] for k in range(len(structs))]
        runs.append((steps, done, res["E"], X, log))
        c.close()
    assert runs[0][1].all()
    assert len(np.unique(runs[0][4][:, 1])) >= 2  # active-structure counts: re-packed at least once
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][2], runs[1][2])
    for a, b in zip(runs[0][3], runs[1][3]):
        assert np.array_equal(a, b)
    # the slowest structures (relaxed through every re-pack) alone
    order = np.argsort(runs[0][0])[::-1][:3]
    for k in order:
        c1 = maise_amd.Context(m, urep=100.0)
        c1.set_batch([structs[k]])
        s1, d1 = c1.relax_lbfgs(max_steps=3000, fmax=0.01)
        assert d1[0] and s1[0] == runs[0][0][k]
        assert c1.batch_result()["E"][0] == pytest.approx(runs[0][2][k], abs=1e-9)
        assert np.abs(c1.result(0, forces=False)["X"] - runs[0][3][k]).max() < 1e-9
        c1.close()
