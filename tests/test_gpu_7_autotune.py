"""GPU tests of the launch-shape autotuner and of the relaxation log.

The autotuner picks 1, 2 or 4 waves per workgroup for the neighbour fill, the descriptor and the force kernel and
chooses stored pair lists or enumeration; none of these choices may change a result bit, so every candidate shape
is checked bit-identical to the default on a batch of uneven structures, and relaxation and MD with the
autotuner switched on (it runs before their first captured iteration) are checked bit-identical to the same runs
without it."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CUPDAG = os.path.join(GOLDEN, "models", "nn_CuPdAg_d3_v1")
CUAG = os.path.join(GOLDEN, "models", "nn_CuAg_d3_v1")


def _batch(n=48, seed=3):
    from maise_amd import workloads as wl
    return wl.cupdag_batch(n, seed=seed)


def test_every_shape_gives_identical_results():
    import maise_amd
    print("[gpu case] autotuner shapes, 48 Cu-Pd-Ag structures", flush=True)
    m = maise_amd.Model(CUPDAG)
    structs = _batch()
    ref = None
    for nbr in (0, 1, 2, 4):
        for wd in (1, 2, 4):
            for store in (True, False):
                c = maise_amd.Context(m, urep=100.0)
                c.set_wpb_nbr(nbr)
                c.set_wpb(wd, wd)
                c.set_pair_store(store)
                c.set_batch(structs)
                c.eval(forces=True)
                r = c.batch_result(forces=True)
                c.close()
                if ref is None:
                    ref = r
                    continue
                np.testing.assert_array_equal(r["E"], ref["E"])
                np.testing.assert_array_equal(r["F"], ref["F"])
                np.testing.assert_array_equal(r["U"], ref["U"])


def test_autotune_picks_a_shape_and_keeps_results():
    import maise_amd
    print("[gpu case] autotune on 48 Cu-Pd-Ag structures", flush=True)
    m = maise_amd.Model(CUPDAG)
    c = maise_amd.Context(m, urep=100.0)
    c.set_batch(_batch())
    c.eval(forces=True)
    r0 = c.batch_result(forces=True)
    shape = c.autotune(reps=2)
    print("tuned shape (nbr, desc, force, pair_store):", shape, flush=True)
    assert shape[0] in (1, 2, 4) and shape[1] in (1, 2, 4) and shape[2] in (1, 2, 4) and shape[3] in (0, 1)
    assert c.shape() == shape
    c.eval(forces=True)
    r1 = c.batch_result(forces=True)
    c.close()
    for k in ("E", "F", "U"):
        np.testing.assert_array_equal(r0[k], r1[k])


def test_relax_with_autotune_is_bit_identical():
    import maise_amd
    print("[gpu case] FIRE with and without the autotuner, 48 Cu-Pd-Ag structures", flush=True)
    m = maise_amd.Model(CUPDAG)
    out = []
    for tune in (False, True):
        c = maise_amd.Context(m, urep=100.0)
        c.set_autotune(tune)
        c.set_batch(_batch())
        steps, done = c.relax(max_steps=400, fmax=0.01)
        r = c.batch_result(forces=True)
        log = c.relax_log()
        out.append((steps, done, r, c.shape()))
        c.close()
        # the log covers the run: iterations grow, active structures never grow, the times are positive
        assert len(log) > 0
        assert np.all(np.diff(log[:, 0]) >= 0) and np.all(np.diff(log[:, 1]) <= 0) and np.all(log[:, 3] > 0)
        assert log[0, 1] == 48
    (s0, d0, r0, sh0), (s1, d1, r1, sh1) = out
    print("shape without / with the autotuner:", sh0, sh1, flush=True)
    assert sh1[1] in (1, 2, 4)
    np.testing.assert_array_equal(s0, s1)
    np.testing.assert_array_equal(d0, d1)
    for k in ("E", "F", "U"):
        np.testing.assert_array_equal(r0[k], r1[k])


def test_md_with_autotune_is_bit_identical():
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] NVT MD with and without the autotuner, Cu-Ag 864 atoms, 40 steps", flush=True)
    s, l, x = wl.alloy_fcc(6, 3.8, 2, 0.0, seed=2)
    m = maise_amd.Model(CUAG)
    out = []
    for tune in (False, True):
        c = maise_amd.Context(m)
        c.set_autotune(tune)
        c.set_structure(s, l, x)
        c.md_init(600.0, seed=1)
        ep, ek = c.md_run(40, dt=1.0, T=600.0, tau=50.0)
        out.append((ep, ek, c.result(0)["X"]))
        c.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])
