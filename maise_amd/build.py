"""In-tree build of libmaise_gpu.so (hipcc for the gfx950 kernels, gcc for the C host code).

Objects go to maise_amd/_build/, the shared library to maise_amd/libmaise_gpu.so and the same objects as a static
archive to maise_amd/libmaise_gpu.a.  Every call rebuilds.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmaise_gpu.so")
STATIC_LIB = os.path.join(HERE, "libmaise_gpu.a")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("MAISE_GPU_ARCH", "gfx950")

HIPCC_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
               # packed FP32 instructions are disabled for all device code
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-command-line-argument"]
CC_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
            f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]

HIP_SOURCES = ["mg_kernels.hip", "mg_mlp.hip", "mg_dyn.hip", "mg_sur.hip", "mg_static.hip", "mg_calib.hip"]
# per-file options.  mg_kernels.hip: MFMA accumulators in ordinary VGPRs (the force kernel reads its results with
# VALU instructions straight away, which the AGPR form would first copy out one register at a time); the MLP
# kernels in mg_mlp.hip measured faster with the AGPR form
HIP_EXTRA = {"mg_kernels.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}
C_SOURCES = ["mg_model.c", "mg_api.c", "mg_dyn.c", "mg_sur.c", "mg_evo.c", "mg_io.c"]
# the search's host operators spread structures over the host cores (OpenMP; libgomp is linked into the library)
C_EXTRA = {"mg_evo.c": ["-fopenmp"]}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", "_hip.o"))
        _run([hipcc, *HIPCC_FLAGS, *HIP_EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    for src in C_SOURCES:
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        _run(["gcc", *CC_FLAGS, *C_EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lgomp", "-lm"], verbose)
    os.replace(tmp, LIB)
    # the same objects as a static archive, for link lines like the reference's (libmaise.a -lm; add the HIP
    # runtime: -L/opt/rocm/lib -lamdhip64 -lstdc++)
    tmpa = STATIC_LIB + ".tmp"
    if os.path.exists(tmpa):
        os.remove(tmpa)
    _run(["ar", "rcs", tmpa, *objs], verbose)
    os.replace(tmpa, STATIC_LIB)
    return LIB


REF_INTERFACE = os.environ.get("MAISE_REFERENCE", "/root/reference") + "/interface"
CALLER_BIN = os.path.join(ROOT, "tests", "drivers", "_bin")


def build_reference_callers(verbose: bool = False) -> list:
    """Compile the reference's own CALL_MAISE callers (interface/c/c2maise.c, interface/cpp/cpp2maise.cpp), unchanged
    from the reference tree, against include/mlib.h and link them to the static library with the reference's link
    line plus the HIP runtime; the executables go to tests/drivers/_bin/ (skipped when the tree is absent)."""
    srcs = [os.path.join(REF_INTERFACE, "c", "c2maise.c"), os.path.join(REF_INTERFACE, "cpp", "cpp2maise.cpp")]
    if not all(os.path.exists(p) for p in srcs):
        return []
    os.makedirs(CALLER_BIN, exist_ok=True)
    out = []
    for src in srcs:
        exe = os.path.join(CALLER_BIN, os.path.splitext(os.path.basename(src))[0])
        _run(["gcc", "-o", exe, src, "-I", os.path.join(ROOT, "include"), f"-I{ROCM}/include", STATIC_LIB,
              f"-L{ROCM}/lib", "-lamdhip64", "-lstdc++", "-lm", "-fopenmp", f"-Wl,-rpath,{ROCM}/lib"], verbose)
        out.append(exe)
    return out


def build_oracle(verbose: bool = False) -> None:
    """Build the CPU oracle and, where the reference tree is present, its reference build (test only)."""
    _run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile"), "all"], verbose)


if __name__ == "__main__":
    build(verbose=True)
    build_oracle(verbose=True)
    print("built", LIB)
