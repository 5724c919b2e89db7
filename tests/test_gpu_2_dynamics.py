"""Device-resident dynamics: NVT/NVE/NPT molecular dynamics and batched FIRE relaxation.

Trajectory parity with the reference is not a bar (its NVT integrator differs and is seeded from the clock,
src/cmod.c:224-453; its minimiser is GSL BFGS2, absent here).  Checked instead: NVE conservation of Epot + Ekin,
the thermostat holding the target temperature, Maxwell-Boltzmann initialisation (zero momentum, exact T), the
barostat's lambda and its drift to the p = 0 volume of the cell relaxation, and
for FIRE that every structure ends below the force threshold with the FP64 oracle agreeing at the relaxed
positions, and that batched results equal one-at-a-time results.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

KB = 1.0 / 11604.0


def _ctx(model):
    import maise_amd
    return maise_amd.Context(maise_amd.Model(model))


def test_md_nve_conserves_energy():
    from maise_amd import workloads as wl
    print("[gpu case] NVE Cu 108 atoms", flush=True)
    c = _ctx(os.path.join(GOLDEN, "models", "nn_Cu_d3_v1"))
    s, l, x = wl.cu_fcc(3, 3.61, 0.02, seed=5)
    c.set_structure(s, l, x)
    c.md_init(300.0, seed=3)
    v = c.md_velocities()
    m = 63.546 * 103.6427
    assert np.abs(v.sum(axis=0)).max() < 1e-12
    assert 0.5 * m * (v ** 2).sum() / (1.5 * len(s) * KB) == pytest.approx(300.0, rel=1e-10)
    ep, ek = c.md_run(300, dt=1.0, T=300.0, tau=0.0)
    tot = ep + ek
    assert np.isfinite(tot).all()
    assert np.abs(tot - tot[0]).max() / len(s) < 2e-4
    c.close()


def test_md_nvt_holds_temperature():
    from maise_amd import workloads as wl
    print("[gpu case] NVT CuAg 256 atoms", flush=True)
    c = _ctx(os.path.join(GOLDEN, "models", "nn_CuAg_d3_v1"))
    s, l, x = wl.alloy_fcc(4, 3.8, 2, 0.02, seed=6)
    c.set_structure(s, l, x)
    c.md_init(600.0, seed=1)
    ep, ek = c.md_run(400, dt=1.0, T=600.0, tau=50.0)
    T = ek / (1.5 * len(s) * KB)
    assert abs(T[200:].mean() - 600.0) < 90.0
    c.close()


def test_md_graph_replay_matches_synchronised_loop():
    """NVT steps replayed from a captured HIP graph, with default headroom and with none (rows outgrow the capture,
    the checkpoint is restored and the block redone), against every step synchronised: the same kernels in the
    same order, so the energy traces, positions and velocities agree exactly."""
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] NVT graph vs synchronised, CuAg 256 atoms 900 K", flush=True)
    m = maise_amd.Model(os.path.join(GOLDEN, "models", "nn_CuAg_d3_v1"))
    s, l, x = wl.alloy_fcc(4, 3.8, 2, 0.02, seed=6)
    out = []
    for enable, headroom in ((True, 0.125), (True, 0.0), (False, 0.125)):
        c = maise_amd.Context(m)
        c.set_graph(enable, headroom)
        c.set_structure(s, l, x)
        c.md_init(900.0, seed=2)
        ep, ek = c.md_run(61, dt=2.0, T=900.0, tau=40.0)
        r = c.result(0)
        out.append((ep, ek, r["X"], r["F"], c.md_velocities()))
        c.close()
    for k in (1, 2):
        for a, b in zip(out[0], out[k]):
            assert np.array_equal(a, b)
    assert np.isfinite(out[0][0]).all()


EV2GPA = 160.2176565


def test_npt_zero_coupling_equals_nvt():
    """With ktt = 0 the barostat factor is exactly 1: the NPT run reproduces the NVT run bit for bit."""
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] NPT ktt=0 vs NVT, Cu 108 atoms", flush=True)
    m = maise_amd.Model(os.path.join(GOLDEN, "models", "nn_Cu_d3_v1"))
    s, l, x = wl.cu_fcc(3, 3.61, 0.02, seed=5)
    out = []
    for npt in (False, True):
        c = maise_amd.Context(m)
        c.set_structure(s, l, x)
        c.md_init(300.0, seed=3)
        if npt:
            ep, ek, vol, pint = c.md_run_npt(40, dt=1.0, T=300.0, tau=50.0, p_gpa=0.0, ktt=0.0)
            assert np.all(vol == vol[0])
        else:
            ep, ek = c.md_run(40, dt=1.0, T=300.0, tau=50.0)
        out.append((ep, ek, c.result(0)["X"], c.lattice()))
        c.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_npt_one_step_lambda():
    """One NPT step scales the lattice by lambda = (1 - (ktt dt/3)(p - p_int))^(1/3) (GPa), with the internal
    pressure p_int = tr(U)/(3V) of the evaluation after the first drift (1e-4 A from the start at 1 K: checked
    against the starting structure's tr(U)/(3V) to 1e-3) in a compressed cell, p_int > 0."""
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] NPT one step, Cu 108 atoms at a = 3.50", flush=True)
    m = maise_amd.Model(os.path.join(GOLDEN, "models", "nn_Cu_d3_v1"))
    s, l, x = wl.cu_fcc(3, 3.50, 0.02, seed=5)
    c = maise_amd.Context(m)
    c.set_structure(s, l, x)
    r0 = c.evaluate()
    V = abs(np.linalg.det(l))
    pint = (r0["U"][0] + r0["U"][1] + r0["U"][2]) / (3.0 * V)
    ktt, dt, p_gpa = 2e-3, 1.0, 1.5
    assert pint * EV2GPA > 5.0  # compressed: the barostat expands
    c.md_init(1.0, seed=2)
    ep, ek, vol, pg = c.md_run_npt(1, dt=dt, T=1.0, tau=0.0, p_gpa=p_gpa, ktt=ktt)
    assert pg[0] == pytest.approx(pint * EV2GPA, rel=1e-3)
    lam = (1.0 - ktt * dt / 3.0 * (p_gpa - pg[0])) ** (1.0 / 3.0)
    assert lam > 1.0
    assert vol[0] == pytest.approx(V * lam ** 3, rel=1e-12)
    np.testing.assert_allclose(c.lattice(), np.asarray(l) * lam, rtol=1e-13, atol=1e-13)
    c.close()


@pytest.mark.parametrize("a0", [3.50, 3.75])
def test_npt_relaxes_toward_equilibrium_volume(a0):
    """Berendsen NPT at p = 0, 300 K from a compressed and an expanded Cu cell: the volume moves toward the p = 0
    volume of the cell relaxation (within 3% over the last third, thermal expansion included) and the internal
    pressure averages near zero there."""
    import maise_amd
    from maise_amd import workloads as wl
    print(f"[gpu case] NPT Cu 108 atoms from a = {a0}", flush=True)
    m = maise_amd.Model(os.path.join(GOLDEN, "models", "nn_Cu_d3_v1"))
    s, l, x = wl.cu_fcc(3, 3.61, 0.0, seed=5)
    cr = maise_amd.Context(m)
    cr.set_structure(s, l, x)
    cr.relax_cell(p_gpa=0.0, max_steps=3000, fmax=1e-3, smax_gpa=1e-3)
    V0 = abs(np.linalg.det(cr.lattice())) / len(s)
    cr.close()
    s, l, x = wl.cu_fcc(3, a0, 0.02, seed=5)
    c = maise_amd.Context(m)
    c.set_structure(s, l, x)
    c.md_init(300.0, seed=4)
    ep, ek, vol, pg = c.md_run_npt(3000, dt=2.0, T=300.0, tau=50.0, p_gpa=0.0, ktt=2e-4)
    v = vol / len(s)
    assert np.isfinite(ep).all() and np.isfinite(v).all()
    tail = v[2000:]
    assert abs(tail.mean() / V0 - 1.0) < 0.03, (tail.mean(), V0)
    assert abs(pg[2000:].mean()) < 1.5, pg[2000:].mean()
    if a0 < 3.6:
        assert tail.mean() > v[0]
    else:
        assert tail.mean() < v[0]
    c.close()


def test_fire_batch_relaxes_and_matches_oracle():
    """Trained ternary potential (random-weight potentials have no minima to relax into), UREP on as in the
    reference executable.  Every structure converges; the FP64 oracle confirms |F| < fmax and the energy at the
    relaxed positions; the oracle's own FIRE (same algorithm, FP64) reaches the same energy per atom."""
    from maise_amd import workloads as wl
    from oracle import pyoracle as po
    path = os.path.join(GOLDEN, "models", "nn_CuPdAg_d3_v1")
    rng = np.random.default_rng(31)
    structs = [wl.random_triclinic(int(rng.integers(8, 33)), 3, rng) for _ in range(12)]
    print("[gpu case] FIRE batch of 12", flush=True)
    import maise_amd
    m = maise_amd.Model(path)
    c = maise_amd.Context(m, urep=100.0)
    c.set_batch(structs)
    steps, done = c.relax(max_steps=2000, fmax=0.01)
    assert done.all(), steps
    res = c.batch_result()
    Xall = c.batch_positions()  # one copy of every structure's positions: the same values as per structure
    om = po.OracleModel(path)
    a = 0
    for k, (spc, lat, _) in enumerate(structs):
        n = len(spc)
        X = c.result(k)["X"]
        np.testing.assert_array_equal(Xall[a:a + n], X)
        a += n
        o = om.evaluate(spc, lat, X, urep=100.0)
        assert np.abs(o["F"]).max() < 0.011
        assert res["E"][k] == pytest.approx(o["E"], abs=1e-5 * n)
    for k in (0, 5):
        spc, lat, x = structs[k]
        _, e_cpu, _, ok = po.fire_relax(om, spc, lat, x, fmax=0.01, urep=100.0)
        assert ok
        assert res["E"][k] / len(spc) == pytest.approx(e_cpu / len(spc), abs=2e-3)
    # one structure relaxed alone gives the same result as inside the batch
    c1 = maise_amd.Context(m, urep=100.0)
    c1.set_batch([structs[3]])
    s1, d1 = c1.relax(max_steps=2000, fmax=0.01)
    assert d1[0] and s1[0] == steps[3]
    assert c1.batch_result()["E"][0] == pytest.approx(res["E"][3], abs=1e-9)
    c.close()
    c1.close()


def test_fire_graph_replay_matches_synchronised_loop():
    """The relaxation replays a captured HIP graph per iteration with capacity-checked edge buffers.  Run the
    same batch three ways: graph with default headroom, graph with zero headroom (rows outgrow the capture
    often, so overflow is detected, the checkpoint restored and the capture redone), and every iteration
    synchronised.  The kernels and their order are the same, so the results agree exactly."""
    import maise_amd
    from maise_amd import workloads as wl
    print("[gpu case] FIRE graph vs synchronised, batch of 10", flush=True)
    path = os.path.join(GOLDEN, "models", "nn_CuPdAg_d3_v1")
    rng = np.random.default_rng(44)
    structs = [wl.random_triclinic(int(rng.integers(8, 40)), 3, rng) for _ in range(10)]
    m = maise_amd.Model(path)
    out = []
    for enable, headroom in ((True, 0.125), (True, 0.0), (False, 0.125)):
        c = maise_amd.Context(m, urep=100.0)
        c.set_graph(enable, headroom)
        c.set_batch(structs)
        steps, done = c.relax(max_steps=1500, fmax=0.01)
        res = c.batch_result()
        X = np.concatenate([c.result(k, forces=False)["X"] for k in range(len(structs))])
        out.append((steps, done, res["E"], res["F"], X))
        c.close()
    for k in (1, 2):
        assert np.array_equal(out[0][0], out[k][0])
        assert np.array_equal(out[0][1], out[k][1])
        assert np.array_equal(out[0][2], out[k][2])
        assert np.array_equal(out[0][3], out[k][3])
        assert np.array_equal(out[0][4], out[k][4])
    assert out[0][1].all()


def test_cell_relax_reaches_reference_minimum():
    """Full-cell relaxation against the reference's relaxed state (path independent, so the end point pins it
    although the reference minimises with GSL BFGS2 and this build with FIRE): test/relax, L1_2 Cu3Pd from
    a = 4.0 A (test/relax/POSCAR) with its model, relaxed to a = 3.72890112684628 A (test/relax/ref/CONTCAR) with
    per-atom energies -4.648947790797 (Cu) and -4.134534446656 eV (Pd) (test/relax/ref/OUTCAR:44-54)."""
    import maise_amd
    print("[gpu case] cell relax test/relax Cu3Pd", flush=True)
    m = maise_amd.Model(os.path.join(GOLDEN, "models", "nn_CuPd_d3_vp_relax"))
    frac = np.array([[0, 0, 0], [0, .5, .5], [.5, .5, 0], [.5, 0, .5]])
    spc = np.array([0, 0, 0, 1], np.int32)
    c = maise_amd.Context(m)
    c.set_structure(spc, np.eye(3) * 4.0, frac, xt=0)
    steps, done = c.relax_cell(p_gpa=0.0, max_steps=4000, fmax=1e-4, smax_gpa=2e-4)
    assert done[0], steps
    lat = c.lattice(0)
    print("relaxed lattice", np.diag(lat), "steps", steps[0], flush=True)
    np.testing.assert_allclose(np.diag(lat), 3.72890112684628, atol=5e-5)
    assert np.abs(lat - np.diag(np.diag(lat))).max() < 1e-6
    r = c.result(0)
    e_ref = 3 * -4.648947790797 + -4.134534446656
    assert abs(r["E"] - e_ref) / 4 < 1e-5  # FP32 path, per atom
    # the relaxed structure itself, evaluated in FP64: the reference's relaxed energies
    from oracle import pyoracle as po
    o = po.OracleModel(os.path.join(GOLDEN, "models", "nn_CuPd_d3_vp_relax")).evaluate(spc, lat, r["X"])
    # (per-atom energies move at first order with the residual displacements; the total is stationary)
    assert o["E"] == pytest.approx(e_ref, abs=2e-7)
    c.close()


def _numbers_close(got, ref, atol, skip=()):
    """Line-by-line: the same tokens where they are not numbers, numbers within atol (per-line override in a
    dict {line: atol}); lines in skip are not compared.  Returns the mismatching lines."""
    bad = []
    for k, (g, r) in enumerate(zip(got, ref)):
        if k in skip:
            continue
        gt, rt = g.split(), r.split()
        ok = len(gt) == len(rt)
        tol = atol.get(k, atol[None]) if isinstance(atol, dict) else atol
        for a, b in zip(gt, rt):
            try:
                ok = ok and abs(float(a) - float(b)) <= tol
            except ValueError:
                ok = ok and a == b
        if not ok:
            bad.append((k, g, r))
    if len(got) != len(ref):
        bad.append(("length", len(got), len(ref)))
    return bad


def test_relax_job_files_match_reference(tmp_path):
    """The file-driven job of test/relax (POSCAR in; OUTCAR, OSZICAR, CONTCAR out) on the device: the CONTCAR has
    the reference's layout line for line (test/relax/ref/CONTCAR, velocity block included) with the relaxed
    lattice within 5e-5 A of the reference's; OUTCAR iteration 0 (the start, a = 4.0) and iteration 1 (the relaxed
    cell) match the reference's blocks 0 and 1 (test/relax/ref/OUTCAR) with the FP32 path's tolerances; the
    "optimizer type" line differs by design (FIRE is 4, the reference ran BFGS2, 0) and the date line is not
    compared."""
    import maise_amd
    from maise_amd import jobs
    print("[gpu case] relax job test/relax files", flush=True)
    files = os.path.join(GOLDEN, "files")
    m = maise_amd.Model(os.path.join(GOLDEN, "models", "nn_CuPd_d3_vp_relax"))
    out = str(tmp_path)
    r = jobs.relax_job(m, os.path.join(files, "relax_POSCAR"), out, rlxt=3, max_steps=4000, fmax=1e-4, smax_gpa=2e-4,
                       version="maise.2.8.17")
    assert r["converged"], r["steps"]
    got = open(os.path.join(out, "CONTCAR")).read().splitlines()
    ref = open(os.path.join(files, "relax_CONTCAR")).read().splitlines()
    assert got[0] == ref[0] and got[5:8] == ref[5:8]
    pos = range(8, 12)  # direct coordinates: compared on the torus (an atom at 0 may come back as 0.99999999...)
    assert _numbers_close(got, ref, 5e-5, skip=pos) == []
    d = np.array([[float(v) for v in got[k].split()] for k in pos]) - \
        np.array([[float(v) for v in ref[k].split()] for k in pos])
    assert np.abs(d - np.round(d)).max() < 1e-5
    got = open(os.path.join(out, "OUTCAR")).read().splitlines()
    ref = open(os.path.join(files, "relax_OUTCAR")).read().splitlines()
    assert len(got) == 58  # header + two iteration blocks
    assert got[3] == "   optimizer type          4" and ref[3] == "   optimizer type          0"
    # FP32 path: atom energies 1e-5 eV, positions/forces 1e-4, lattice 5e-5 A, stress 1e-6 eV/A^3 (kB: x1602)
    tol = {None: 1e-4, 28: 2e-6, 29: 4e-3, 31: 5e-5, 39: 1e-3, 50: 2e-6, 51: 4e-3, 53: 5e-5}
    pos = list(range(23, 27)) + list(range(45, 49))  # position rows: Cartesian columns compared on the torus
    assert _numbers_close(got, ref[:58], tol, skip=[1, 3] + pos) == []
    for k in pos:
        g, q = np.array(got[k].split(), float), np.array(ref[k].split(), float)
        a = 4.0 if k < 40 else 3.72890112684628
        d = (g[:3] - q[:3]) / a
        assert np.abs(d - np.round(d)).max() * a < 1e-4, (k, got[k], ref[k])
        np.testing.assert_allclose(g[3:6], q[3:6], atol=1e-4)
        assert abs(g[6] - q[6]) < 1e-5, (k, got[k], ref[k])
    oz = open(os.path.join(out, "OSZICAR")).read().split()
    np.testing.assert_allclose([float(v) for v in oz], [-17.25531076, -18.08137782], atol=5e-5)


def test_cell_relax_examples_03rlx_final_energy():
    """examples/03-rlx: Cu4 in a 3.626 x 3.626 x 6.026 A cell with its model relaxes (positions + cell) to the
    minimum H = -16.3965601100172194 eV (examples/03-rlx/README:48)."""
    import maise_amd
    from maise_amd import poscar
    print("[gpu case] cell relax examples/03-rlx", flush=True)
    p = poscar.read_poscar(os.path.join(GOLDEN, "structures", "POSCAR_03rlx"))
    c = maise_amd.Context(maise_amd.Model(os.path.join(GOLDEN, "models", "nn_Cu_03rlx")))
    c.set_structure(p.species, p.lattice, p.frac, xt=0)
    steps, done = c.relax_cell(p_gpa=0.0, max_steps=6000, fmax=1e-4, smax_gpa=2e-4)
    assert done[0], steps
    E = c.result(0)["E"]
    print("03-rlx final E", E, "steps", steps[0], "lattice", c.lattice(0).round(6).tolist(), flush=True)
    assert E == pytest.approx(-16.3965601100172194, abs=4e-5)  # FP32 path: 1e-5 eV/atom
    # the relaxed structure itself, evaluated in FP64, sits at the reference's minimum
    from oracle import pyoracle as po
    o = po.OracleModel(os.path.join(GOLDEN, "models", "nn_Cu_03rlx")).evaluate(p.species, c.lattice(0),
                                                                               c.result(0)["X"])
    assert o["E"] == pytest.approx(-16.3965601100172194, abs=1e-7)
    c.close()
