"""GPU tests of the skin neighbour lists: MD and FIRE with candidates kept within Rc + skin (rebuilt only after an
atom has moved skin/2) give bit-identical trajectories to the same runs that rebuild the lists every step, on a
large periodic cell (Cu-Ag alloy, many bins), on uneven triclinic cells (Cu-Pd-Ag batch) and on small cells whose
periodic images are many (Mg16Ca8, one bin per cell), where atoms cross cell faces and their image shifts change."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CUAG = os.path.join(GOLDEN, "models", "nn_CuAg_d3_v1")
CUPDAG = os.path.join(GOLDEN, "models", "nn_CuPdAg_d3_v1")
MGCA = os.path.join(GOLDEN, "models", "nn_MgCa_d3_v0")


def test_md_with_skin_lists_is_bit_identical():
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] NVT MD 864-atom Cu-Ag, 300 steps, skin 0 / 0.3 A", flush=True)
    s, l, x = wl.alloy_fcc(6, 3.8, 2, 0.0, seed=2)
    m = maise_amd.Model(CUAG)
    out = []
    for skin in (0.0, 0.3):
        c = maise_amd.Context(m)
        c.set_skin(skin)
        c.set_structure(s, l, x)
        c.md_init(600.0, seed=1)
        ep, ek = c.md_run(300, dt=1.0, T=600.0, tau=50.0)
        out.append((ep, ek, c.result(0)["X"], c.skin_builds()))
        c.close()
    print("skin builds:", out[1][3], flush=True)
    assert 2 <= out[1][3] < 300
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])


def test_md_twice_on_one_context_with_skin_lists():
    """A second run on the same context starts from the first run's capacities and candidate rows: re-sizing them
    before the capture must mark the rows stale (regression: a stride change left rows of the old stride in use)."""
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] NVT MD 864-atom Cu-Ag, 20 + 100 steps on one context, skin 0 / 0.5 A", flush=True)
    s, l, x = wl.alloy_fcc(6, 3.8, 2, 0.0, seed=2)
    m = maise_amd.Model(CUAG)
    out = []
    for skin in (0.0, 0.5):
        c = maise_amd.Context(m)
        c.set_skin(skin)
        c.set_structure(s, l, x)
        c.md_init(600.0, seed=1)
        c.md_run(20, dt=1.0, T=600.0, tau=50.0)
        c.md_init(600.0, seed=3)
        ep, ek = c.md_run(100, dt=1.0, T=600.0, tau=50.0)
        out.append((ep, ek, c.result(0)["X"]))
        c.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def _relax_pair(model, structs, skin, urep=100.0, max_steps=600):
    import maise_amd
    m = maise_amd.Model(model)
    res = []
    for sk in (0.0, skin):
        c = maise_amd.Context(m, urep=urep)
        c.set_skin(sk)
        c.set_batch(structs)
        steps, done = c.relax(max_steps=max_steps, fmax=0.01)
        r = c.batch_result(forces=True)
        res.append((steps, done, r, c.skin_builds()))
        c.close()
    (s0, d0, r0, _), (s1, d1, r1, b1) = res
    np.testing.assert_array_equal(s0, s1)
    np.testing.assert_array_equal(d0, d1)
    for k in ("E", "F", "U"):
        np.testing.assert_array_equal(r0[k], r1[k])
    return b1


def test_fire_with_skin_lists_is_bit_identical_triclinic_batch():
    from maise_amd import workloads as wl
    print("[gpu case] FIRE 48 Cu-Pd-Ag structures, skin 0 / 0.5 A", flush=True)
    b = _relax_pair(CUPDAG, wl.cupdag_batch(48, seed=3), 0.5)
    print("skin builds:", b, flush=True)
    assert b >= 1


def test_fire_with_skin_lists_is_bit_identical_small_cells():
    from maise_amd.evolve import GAConfig, random_structure
    print("[gpu case] FIRE 32 Mg16Ca8 cells, skin 0 / 0.5 A", flush=True)
    cfg = GAConfig()
    rng = np.random.default_rng(7)
    structs = [random_structure(cfg, rng) for _ in range(32)]
    b = _relax_pair(MGCA, structs, 0.5)
    print("skin builds:", b, flush=True)
    assert b >= 1


def test_cell_changing_runs_ignore_the_skin_setting():
    """NPT MD and cell relaxation move the lattice every step, so they rebuild the lists every step whatever the
    skin: with a skin set they reproduce the runs without one bit for bit and build no candidates."""
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] NPT MD and cell relaxation with skin 0 / 1.0 A, Cu 108 atoms", flush=True)
    m = maise_amd.Model(os.path.join(GOLDEN, "models", "nn_Cu_d3_v1"))
    s, l, x = wl.cu_fcc(3, 3.55, 0.02, seed=5)
    out = []
    for skin in (0.0, 1.0):
        c = maise_amd.Context(m)
        c.set_skin(skin)
        c.set_structure(s, l, x)
        c.md_init(300.0, seed=3)
        ep, ek, vol, pint = c.md_run_npt(60, dt=1.0, T=300.0, tau=50.0, p_gpa=0.0, ktt=2e-3)
        c.set_structure(s, l, x)
        steps, done = c.relax_cell(p_gpa=0.0, max_steps=200, fmax=0.01, smax_gpa=0.05)
        out.append((ep, ek, vol, pint, steps, c.lattice(), c.result(0)["X"], c.skin_builds()))
        c.close()
    for a, b in zip(out[0][:7], out[1][:7]):
        np.testing.assert_array_equal(a, b)
    assert out[1][7] == 0
