"""GPU parity for kernel variants the reference-build golden cases do not reach, against the FP64 oracle.

* generic (eta, zeta, lambda) grid with every species pair sharing its G4 etas and only two zetas {1, 3}: the
  shared-eta matrix-core force contraction with unused zeta slots (their 2^(1 - zeta) prefactor slot is 0);
* the 43-function table with G4 etas up to 2.0 A^-2: radial factors exp(-eta (r_ij^2 + r_ik^2 + r_jk^2)) down to
  ~1e-30, far below the FP16 normal range the split matrix-core operands use;
* hidden layers wider than 64 (51-96-96-1), which run on the FP32 VALU MLP kernel instead of MFMA;
* hidden layers of one and two 32-unit blocks mixed (64-10 on the binary alloy, 10-48 on Cu): the two-wave MFMA
  MLP kernel's layouts where one layer is split by output block between the waves and the other is not.

Each synthetic model has seeded weights and input ranges set from the oracle's descriptors of the test
structure, so the network is neither saturated nor trivially linear.  Bars are the FP32 path's: energy 1e-5
eV/atom, forces 1e-4 eV/A, virial 1e-4 eV/atom.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

E_TOL, F_TOL, U_TOL = 1e-5, 1e-4, 1e-4


def _structures(kind):
    from maise_amd import workloads as wl
    if kind == "cu":
        return wl.cu_fcc(4, 3.61, 0.08, seed=11)
    spc, lat, x = wl.alloy_fcc(3, 3.8, nspc=2, sigma=0.08, seed=12)
    return spc, lat, x


def _model(tmp, name, spcz, hidden, basis):
    from maise_amd import workloads as wl
    from oracle import pyoracle as po
    path = str(tmp / name)
    kind = "cu" if len(spcz) == 1 else "cuag"
    sample = [_structures(kind)]

    def desc(p, s, l, x):
        return po.OracleModel(p).evaluate(s, l, x, want_g=True, forces=False)["G"]

    wl.calibrated_model(path, spcz, hidden, seed=21, sample_structures=sample, descriptor_fn=desc, basis=basis)
    return path, kind


def _variants():
    from maise_amd import modelfile as mf
    return {
        "generic_shared_z13_cu": ([29], (10, 10), mf.generic_basis(zetas=(1.0, 3.0))),
        "generic_shared_z13_cuag": ([29, 47], (10, 10), mf.generic_basis(zetas=(1.0, 3.0))),
        "large_eta_cu": ([29], (10, 10), mf.standard_basis(eta4=np.geomspace(0.05, 2.0, 7))),
        "wide_mlp_cu": ([29], (96, 96), None),
        "mlp_64_10_cuag": ([29, 47], (64, 10), None),
        "mlp_10_48_cu": ([29], (10, 48), None),
    }


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("variants")
    return {name: _model(tmp, name, *spec) for name, spec in _variants().items()}


@pytest.mark.parametrize("name", ["generic_shared_z13_cu", "generic_shared_z13_cuag", "large_eta_cu", "wide_mlp_cu",
                                  "mlp_64_10_cuag", "mlp_10_48_cu"])
def test_variant_matches_oracle(built, name):
    import maise_amd
    from oracle import pyoracle as po
    path, kind = built[name]
    spc, lat, x = _structures(kind)
    m = maise_amd.Model(path)
    print(f"[gpu case] variant {name}: N={len(spc)} D={m.D} H1={m.H1} g4={m.g4_variant} shared_eta={m.shared_eta} "
          f"mfma={m.mfma}", flush=True)
    if name.startswith("generic"):
        assert m.g4_variant == 0 and m.shared_eta
    if name == "large_eta_cu":
        assert m.g4_variant == 1 and m.shared_eta
    if name == "wide_mlp_cu":
        assert not m.mfma
    if name.startswith("mlp_"):
        assert m.mfma
    c = maise_amd.Context(m)
    c.set_structure(spc, lat, x)
    r = c.evaluate()
    c.close()
    o = po.OracleModel(path).evaluate(spc, lat, x, threshold=False)
    N = len(spc)
    de = abs(r["E"] - o["E"]) / N
    df = np.abs(r["F"] - o["F"]).max()
    du = np.abs(r["U"] - o["U"]).max() / N
    fmax = np.abs(o["F"]).max()
    print(f"{name}: dE/atom={de:.2e} dF={df:.2e} (|F|max {fmax:.2e}) dU/atom={du:.2e}")
    assert np.isfinite(r["F"]).all() and np.isfinite(r["U"]).all()
    assert fmax > 1e-2  # the structure is jittered: forces are not trivially zero
    assert de < E_TOL
    assert df < F_TOL
    assert du < U_TOL
