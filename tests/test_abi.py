"""The C ABI: libmaise_gpu.so loads (no GPU needed) and exports every entry point include/*.h declares; the
drop-in header gives the reference's struct sizes and field offsets (measured on the reference headers with
the same compiler: ANN 4704, PRS 160, LNK 376, Cell 10624, LNK.B at 152, Cell.p at 1736); a caller written like
interface/c/c2maise.c compiles and links against it unchanged."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "maise_amd", "libmaise_gpu.so")


def declared_functions():
    names = set()
    for h in ("maise_gpu.h", "mlib.h"):
        txt = open(os.path.join(INC, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "sizeof", "static_assert", "_Static_assert", "MG_STATIC_ASSERT"):
                names.add(name)
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert "CALL_MAISE" in names and "mg_eval" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_dropin_struct_layout(tmp_path):
    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-I", INC, "-I/opt/rocm/include", "-o", str(exe),
                    os.path.join(ROOT, "tests", "drivers", "abi_layout.c")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert [int(v) for v in out] == [4704, 160, 376, 10624, 152, 1736, 8]


def test_reference_style_caller_links(tmp_path):
    exe = tmp_path / "drv"
    subprocess.run(["gcc", "-O2", "-I", INC, "-I/opt/rocm/include", "-o", str(exe),
                    os.path.join(ROOT, "tests", "drivers", "call_maise_driver.c"), LIB,
                    "-Wl,-rpath," + os.path.dirname(LIB), "-lm"], check=True)
    assert exe.exists()


def test_model_reader_matches_oracle_reader():
    """Host-side model parsing needs no GPU: sizes, cutoff and species agree with the oracle's reader."""
    import maise_amd
    from oracle import pyoracle as po
    for name in ("nn_CuAg_d3_v0_iface", "nn_CuPdAg_d3_v1", "nn_CuPd_d3_vp_relax"):
        p = os.path.join(ROOT, "tests", "golden", "models", name)
        a, b = maise_amd.Model(p), po.OracleModel(p)
        assert (a.nspc, a.D, a.H1, a.H2, a.nsym, a.ng2, a.ng4, a.nw) == (b.nspc, b.D, b.H1, b.H2, b.nsym, b.ng2, b.ng4, b.nw)
        assert a.spcz == b.spcz and a.rc == b.rc


def test_model_reader_rejects_bad_files(tmp_path):
    import maise_amd
    p = tmp_path / "model"
    p.write_text("not a model\n")
    with pytest.raises(maise_amd.MaiseError):
        maise_amd.Model(str(p))


def test_static_eval_rejects_oversize_before_any_gpu_work():
    """The evaluation that allocates nothing refuses structures beyond its fixed scratch with a clear message
    (checked before any device call, so this runs without a GPU)."""
    import numpy as np
    import maise_amd
    m = maise_amd.Model(os.path.join(ROOT, "maise_amd", "potentials", "cu_51_10_10.nn"))
    N = 32769
    with pytest.raises(maise_amd.MaiseError, match="fixed-scratch limit"):
        maise_amd.eval_static(m, np.zeros(N, np.int32), np.eye(3) * 100.0, np.zeros((N, 3)))


REF_IFACE = os.path.join(os.environ.get("MAISE_REFERENCE", "/root/reference"), "interface")


@pytest.mark.skipif(not os.path.isdir(REF_IFACE), reason="reference tree not mounted")
@pytest.mark.parametrize("lib_kind", ["static", "shared"])
def test_reference_callers_compile_unchanged(tmp_path, lib_kind):
    """The reference's own callers, interface/c/c2maise.c and interface/cpp/cpp2maise.cpp, read from the reference
    tree and compiled unchanged against include/mlib.h, link against libmaise_gpu.a with the reference's link line
    (libmaise.a -lm -fopenmp, interface/c/makefile:6) plus the HIP runtime, and against libmaise_gpu.so."""
    lib = os.path.join(ROOT, "maise_amd", "libmaise_gpu.a" if lib_kind == "static" else "libmaise_gpu.so")
    assert os.path.exists(lib)
    for src in (os.path.join(REF_IFACE, "c", "c2maise.c"), os.path.join(REF_IFACE, "cpp", "cpp2maise.cpp")):
        exe = tmp_path / (os.path.basename(src) + "." + lib_kind)
        cmd = ["gcc", "-o", str(exe), src, "-I", INC, "-I/opt/rocm/include", lib, "-L/opt/rocm/lib", "-lamdhip64",
               "-lstdc++", "-lm", "-fopenmp", "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath," + os.path.dirname(lib)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert exe.exists()
