"""GPU parity: the HIP path through the C ABI against the reference build's golden vectors and the oracle.

Bars: neighbour sets and wrapped positions bit-exact; descriptors within FP32 rounding; energy within
1e-5 eV/atom, forces within 1e-4 eV/A, virial within 1e-4 eV/atom (FP32 path vs FP64).  At the full 32,000-atom
size, size-independent properties: supercell additivity against the oracle's 256-atom cell, zero forces on the
perfect lattice, translation invariance and sum F = 0.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, model_path, nbr_sets

pytestmark = pytest.mark.gpu

# kernel variants covered: 43-function table (cu, cuag, mgca, config1, ...), 22-function table (relax), generic
# grid with per-pair etas (species_generic), 43-function table with per-pair etas (species_std)
CASES = ["iface", "relax", "cu", "cuag", "cupdag_tri", "urep", "small_tri", "mgca", "config1", "config1_w64",
         "shell", "config3_like", "species_generic", "species_std"]

E_TOL = 1e-5   # eV/atom
F_TOL = 1e-4   # eV/A
U_TOL = 1e-4   # eV/atom


def gpu_eval(case):
    import maise_amd
    print(f"[gpu case] model={case['model']} N={len(case['spc'])} nd={int(case['nd'])}", flush=True)
    m = maise_amd.Model(model_path(case["model"]))
    c = maise_amd.Context(m, urep=float(case["urep"]))
    c.set_structure(case["spc"], case["lat"], case["x"], nd=int(case["nd"]), xt=int(case["xt"]))
    c.eval(forces=True, keep_g=True)
    r = c.result()
    r["G"] = c.descriptors()
    r["row"], r["j"], r["img"] = c.neighbors()
    c.close()
    return r


@pytest.fixture(scope="module")
def results(ref_cases):
    return {name: gpu_eval(ref_cases[name]) for name in CASES}


@pytest.mark.parametrize("name", CASES)
def test_neighbour_sets_bit_exact(ref_cases, results, name):
    c, r = ref_cases[name], results[name]
    counts = np.diff(r["row"])
    np.testing.assert_array_equal(counts, c["nn"])
    ent = np.concatenate([r["j"][:, None], r["img"]], axis=1)
    assert nbr_sets(counts, ent) == nbr_sets(c["nn"], c["nbr"])


@pytest.mark.parametrize("name", CASES)
def test_wrapped_positions_bit_exact(ref_cases, results, name):
    np.testing.assert_array_equal(results[name]["X"], ref_cases[name]["X"])


@pytest.mark.parametrize("name", [n for n in CASES if n != "config1_w64"])
def test_descriptors(ref_cases, results, name):
    G, Gr = results[name]["G"].astype(np.float64), ref_cases[name]["G"]
    err = np.abs(G - Gr) / (1.0 + np.abs(Gr))
    assert err.max() < 2e-5, err.max()


def exact_gradient(case):
    """FP64 -dE/dx and virial from the oracle.  For models whose basis blocks differ per species pair the
    reference's derivative code (src/nprs.c:597-738) evaluates every pair with the species-0 block, so its forces
    are not the gradient of its own energy; the GPU path computes the gradient, and is held to it."""
    from oracle import pyoracle as po
    o = po.OracleModel(model_path(case["model"])).evaluate(case["spc"], case["lat"], case["x"], nd=int(case["nd"]),
                                                          xt=int(case["xt"]), urep=float(case["urep"]),
                                                          threshold=False)
    return o["F"], o["U"]


@pytest.mark.parametrize("name", CASES)
def test_energy_forces_virial(ref_cases, results, name):
    c, r = ref_cases[name], results[name]
    N = len(c["spc"])
    Fr, Ur = exact_gradient(c) if name.startswith("species_") else (c["F"], c["U"])
    de = abs(r["E"] - float(c["E"])) / N
    dea = np.abs(r["EA"] - c["EA"]).max()
    df = np.abs(r["F"] - Fr).max()
    du = np.abs(r["U"] - Ur).max() / N
    print(f"{name}: dE/atom={de:.2e} dEA={dea:.2e} dF={df:.2e} dU/atom={du:.2e}")
    assert de < E_TOL
    assert dea < 10 * E_TOL
    assert df < F_TOL
    assert du < U_TOL


def test_call_maise_dropin(ref_cases, tmp_path):
    """A caller written like interface/c/c2maise.c, linked to libmaise_gpu.so, reproduces the reference's
    CALL_MAISE outputs (H incl. pV, thresholded FRC/STR) on the interface cluster and the relax fixture."""
    exe = tmp_path / "drv"
    inc = os.path.join(ROOT, "include")
    lib = os.path.join(ROOT, "maise_amd", "libmaise_gpu.so")
    subprocess.run(["gcc", "-O2", "-I", inc, "-I/opt/rocm/include", "-o", str(exe),
                    os.path.join(ROOT, "tests", "drivers", "call_maise_driver.c"), lib,
                    "-Wl,-rpath," + os.path.dirname(lib), "-lm"], check=True)
    for name in ("iface", "relax_p"):
        c = ref_cases[name]
        d = tmp_path / name
        d.mkdir()
        N = len(c["spc"])
        with open(d / "input", "w") as f:
            f.write(f"{N} {int(c['nd'])} {int(c['xt'])} 500 {float(c['p_gpa'])!r}\n")
            f.write(" ".join(repr(float(v)) for v in c["lat"].ravel()) + "\n")
            for i in range(N):
                f.write(f"{int(c['spc'][i])} " + " ".join(repr(float(v)) for v in c["x"][i]) + "\n")
        env = dict(os.environ, MAISE_MODEL=model_path(c["model"]), MAISE_UREP=repr(float(c["urep"])))
        print(f"[gpu case] CALL_MAISE {name}", flush=True)
        out = subprocess.run([str(exe), "input", "3"], cwd=d, env=env, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        lines = out.stdout.split("\n")
        H = float(lines[0].split()[1])
        F = np.array([[float(v) for v in l.split()[1:]] for l in lines if l.startswith("F ")])
        S = np.array([float(l.split()[1]) for l in lines if l.startswith("S ")])
        assert H == pytest.approx(float(c["H"]), abs=E_TOL * N)
        np.testing.assert_allclose(F, c["F"], atol=F_TOL)
        np.testing.assert_allclose(S, c["U"], atol=U_TOL * N)


@pytest.mark.parametrize("exe", ["c2maise", "cpp2maise"])
def test_reference_callers_run_unchanged(ref_cases, tmp_path, exe):
    """The reference's own interface drivers (interface/c/c2maise.c, interface/cpp/cpp2maise.cpp), compiled unchanged
    against include/mlib.h and linked to libmaise_gpu.a by build() where the reference tree is present, run on the
    GPU with the interface's model (a copy: tests/golden/models/nn_CuAg_d3_v0_iface) in their working directory and
    print the forces and enthalpy of the reference's interface/c/test/OUTCAR (4-atom Cu cluster)."""
    path = os.path.join(ROOT, "tests", "drivers", "_bin", exe)
    if not os.path.exists(path):
        pytest.skip("reference callers not built (the reference tree was absent at build time)")
    c = ref_cases["iface"]
    with open(os.path.join(GOLDEN, "models", "nn_CuAg_d3_v0_iface")) as src, open(tmp_path / "model", "w") as dst:
        dst.write(src.read())
    print(f"[gpu case] reference caller {exe}", flush=True)
    env = {k: v for k, v in os.environ.items() if k not in ("MAISE_MODEL", "MAISE_UREP")}
    out = subprocess.run([path], cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    F = np.array([[float(v) for v in l.split()[1:4]] for l in out.stdout.splitlines()
                  if len(l.split()) == 4 and l.split()[0].isdigit()])
    H = float(next(l.split()[1] for l in out.stdout.splitlines() if l.startswith("Enthalpy")))
    print(out.stdout)
    assert F.shape == (4, 3)
    assert H == pytest.approx(float(c["H"]), abs=E_TOL * 4)
    np.testing.assert_allclose(F, c["F"], atol=F_TOL)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("name", ["cuag", "mgca", "cupdag_tri", "relax", "species_generic", "species_std", "config1"])
def test_stored_pairs_match_enumeration(ref_cases, name, prec):
    """The force kernel reading the descriptor kernel's stored pair lists visits the same pairs in the same order
    as enumerating them itself: bit-identical energies, forces and virials (43- and 22-function tables, generic
    grid, shared and per-pair etas, 1-3 species), with the VALU contraction (FP32) and with the matrix-core one
    (compensated BF16: a triplet's terms must not depend on the other triplets of its batch)."""
    import maise_amd
    case = ref_cases[name]
    precision = maise_amd.MG_PREC_FP32 if prec == "fp32" else maise_amd.MG_PREC_BF16X3
    out = []
    for store in (True, False):
        print(f"[gpu case] {name} {prec} pair_store={store}", flush=True)
        m = maise_amd.Model(model_path(case["model"]))
        c = maise_amd.Context(m, urep=float(case["urep"]), precision=precision)
        c.set_pair_store(store)
        c.set_structure(case["spc"], case["lat"], case["x"], nd=int(case["nd"]), xt=int(case["xt"]))
        c.eval(forces=True)
        r1 = c.result()
        c.eval(forces=True)  # the second evaluation runs with the capacities of the first (no host sync inside)
        r2 = c.result()
        c.close()
        for k in ("F", "U"):
            np.testing.assert_array_equal(r1[k], r2[k])
        out.append(r1)
    assert out[0]["E"] == out[1]["E"]
    np.testing.assert_array_equal(out[0]["F"], out[1]["F"])
    np.testing.assert_array_equal(out[0]["U"], out[1]["U"])


def test_full_size_32k_properties():
    """Cu fcc 20x20x20 (32,000 atoms), synthetic 51-64-64-1 at Rc 6.0: per-atom energy of the perfect lattice
    equals the oracle's 256-atom cell (supercell additivity), forces vanish by symmetry, 78 neighbours each;
    jittered: translation invariance and sum F = 0."""
    import maise_amd
    from maise_amd import workloads as wl
    from oracle import pyoracle as po
    path = os.path.join(ROOT, "maise_amd", "potentials", "cu_51_64_64.nn")
    s1, l1, x1 = wl.cu_fcc(4, 3.61, 0.0)
    e_ref = po.OracleModel(path).evaluate(s1, l1, x1, forces=False)["E"] / len(s1)
    m = maise_amd.Model(path)
    c = maise_amd.Context(m)
    s, l, x = wl.cu_fcc(20, 3.61, 0.0)
    print("[gpu case] 32k perfect", flush=True)
    c.set_structure(s, l, x)
    r = c.evaluate()
    row, _, _ = c.neighbors()
    assert np.all(np.diff(row) == 78)
    assert abs(r["E"] / len(s) - e_ref) < E_TOL
    assert np.abs(r["F"]).max() < F_TOL
    s, l, x = wl.cu_fcc(20, 3.61, 0.05, seed=4)
    print("[gpu case] 32k jittered", flush=True)
    c.set_structure(s, l, x)
    a = c.evaluate()
    c.set_structure(s, l, x + np.array([0.3, -0.7, 1.1]))
    b = c.evaluate()
    assert abs(a["E"] - b["E"]) / len(s) < E_TOL
    assert np.abs(a["F"] - b["F"]).max() < F_TOL
    assert np.abs(a["F"].sum(axis=0)).max() < 1e-2
    c.close()


@pytest.mark.parametrize("name", ["config1", "config1_w64", "cuag", "cupdag_tri", "mgca"])
def test_bf16_mlp_error_bounds(ref_cases, name):
    """Compensated BF16 MLP on MFMA (split hi/lo operands, FP32 accumulation, FP32 descriptors): energy within
    1 meV/atom and forces within 10 meV/A of the FP64 reference build (measured ~1e-6 eV/atom).  The one-pass BF16
    mode is only reported, not bounded: its weight rounding alone gives ~2e-3 eV/atom on the 51-10-10-1 model
    (emulated on the CPU with the same rounding) and up to ~2e-2 eV/atom, 0.3 eV/A on the shipped models."""
    import maise_amd
    c = ref_cases[name]
    print(f"[gpu case] bf16 {name}", flush=True)
    m = maise_amd.Model(model_path(c["model"]))
    N = len(c["spc"])
    for prec, etol, ftol in ((maise_amd.MG_PREC_BF16X3, 1e-3, 1e-2), (maise_amd.MG_PREC_BF16, np.inf, np.inf)):
        ctx = maise_amd.Context(m, urep=float(c["urep"]), precision=prec)
        ctx.set_structure(c["spc"], c["lat"], c["x"], nd=int(c["nd"]), xt=int(c["xt"]))
        r = ctx.evaluate()
        de = abs(r["E"] - float(c["E"])) / N
        df = np.abs(r["F"] - c["F"]).max()
        print(f"bf16 prec={prec} {name}: dE/atom={de:.2e} dF={df:.2e}")
        assert de < etol
        assert df < ftol
        ctx.close()


@pytest.mark.parametrize("a", [2.4, 1.9])
def test_long_rows_match_oracle(a):
    """Compressed Cu fcc (2x2x2 cells, jittered) with rows of about 270 and 540 neighbours: LDS tiles of 512 and
    1,024 entries, two and one wave per block in the neighbour fill, reverse-lookup tables of 1,024 and 2,048 slots.
    Neighbour sets bit-exact against the oracle's LIST restatement; descriptors, energy and forces against the
    oracle in FP64 (FP32 sums over several hundred neighbours: relative bars)."""
    import maise_amd
    from oracle import pyoracle as po
    print(f"[gpu case] long rows Cu fcc a={a}", flush=True)
    rng = np.random.default_rng(17)
    cell = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    frac = np.concatenate([(cell + np.array([i, j, k])) / 2.0 for i in range(2) for j in range(2) for k in range(2)])
    lat = np.eye(3) * 2 * a
    x = frac @ lat + rng.normal(0, 0.02, frac.shape)
    spc = np.zeros(len(x), np.int32)
    path = os.path.join(GOLDEN, "models", "nn_Cu_d3_v1")
    m = maise_amd.Model(path)
    c = maise_amd.Context(m)
    c.set_structure(spc, lat, x)
    c.eval(forces=True, keep_g=True)
    r = c.result()
    G = c.descriptors().astype(np.float64)
    row, j, img = c.neighbors()
    c.close()
    om = po.OracleModel(path)
    cnt, ent, _ = po.neighbors(lat, x, om.rc)
    counts = np.diff(row)
    print("row lengths", counts.min(), counts.max(), flush=True)
    assert counts.max() > (250 if a > 2 else 500)
    np.testing.assert_array_equal(counts, cnt)
    assert nbr_sets(counts, np.concatenate([j[:, None], img], axis=1)) == nbr_sets(cnt, ent)
    o = om.evaluate(spc, lat, x, want_g=True, threshold=False)
    assert (np.abs(G - o["G"]) / (1.0 + np.abs(o["G"]))).max() < 1e-4
    N = len(x)
    assert abs(r["E"] - o["E"]) / N < 1e-4 * (1.0 + abs(o["E"]) / N)
    assert np.abs(r["F"] - o["F"]).max() < 1e-4 * (1.0 + np.abs(o["F"]).max())
