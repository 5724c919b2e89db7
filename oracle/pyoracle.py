"""ctypes access to the FP64 CPU oracle (liboracle.so) and to the reference build (oracle/_ref/maise_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by bench.py's CPU-baseline leg.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_EXE = os.path.join(HERE, "_ref", "maise_ref")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -f oracle/Makefile)")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.orc_load.argtypes = [C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_int]
        L.orc_load.restype = C.c_int
        L.orc_free.argtypes = [vp]
        L.orc_info.argtypes = [vp, ip, dp]
        L.orc_eval.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, dp, dp, C.c_double, C.c_double, C.c_int,
                               C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp, dp, ip]
        L.orc_eval.restype = C.c_int
        L.orc_neighbors.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, C.c_double, C.c_int, ip, ip,
                                    C.c_longlong, dp]
        L.orc_neighbors.restype = C.c_int
        L.orc_mlp.argtypes = [vp, C.c_int, dp, dp]
        L.orc_mlp.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        L.orc_jar.argtypes = [C.c_int, dp, dp]
        L.orc_set_ref_derivatives.argtypes = [C.c_int]
        L.orc_set_ref_derivatives.restype = None
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


class OracleModel:
    def __init__(self, path: str):
        self._h = C.c_void_p()
        err = C.create_string_buffer(256)
        if lib().orc_load(path.encode(), C.byref(self._h), err, 256) != 0:
            raise RuntimeError(f"oracle cannot load {path}: {err.value.decode()}")
        info = np.zeros(12, dtype=np.int32)
        rc = C.c_double()
        lib().orc_info(self._h, _ip(info), C.byref(rc))
        self.nspc, self.D, self.H1, self.H2, self.nsym, self.ng2, self.ng4, self.nw = (int(v) for v in info[:8])
        self.spcz = [int(v) for v in info[8:8 + self.nspc]]
        self.rc = rc.value

    def __del__(self):
        try:
            if self._h:
                lib().orc_free(self._h)
        except Exception:
            pass

    def evaluate(self, spc, lat, x, nd=3, xt=1, p=0.0, urep=0.0, forces=True, threshold=True, cap=4096,
                 want_g=False, ref_derivatives=False):
        """FP64 E/F/S of one structure.  p is in eV/A^3 (CALL_MAISE divides its GPa input by eV2GPa).
        ref_derivatives: forces and virial as the reference's derivative code computes them (species-0 basis
        parameters for every pair; differs from -dE/dx only for models with species-specific basis blocks)."""
        spc = np.ascontiguousarray(spc, dtype=np.int32)
        lat = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        N = spc.size
        H, E = C.c_double(), C.c_double()
        EA = np.zeros(N)
        F = np.zeros(3 * N)
        U = np.zeros(6)
        G = np.zeros(N * self.D) if want_g else None
        g = np.zeros(N * self.D) if want_g else None
        Xw = np.zeros(3 * N)
        nn = np.zeros(N, dtype=np.int32)
        lib().orc_set_ref_derivatives(1 if ref_derivatives else 0)
        try:
            rc = lib().orc_eval(self._h, N, nd, xt, _ip(spc), _dp(lat), _dp(x), p, urep, 1 if forces else 0,
                                1 if threshold else 0, cap, C.byref(H), C.byref(E), _dp(EA), _dp(F), _dp(U), _dp(G),
                                _dp(g), _dp(Xw), _ip(nn))
        finally:
            lib().orc_set_ref_derivatives(0)
        if rc != 0:
            raise RuntimeError("oracle: neighbour capacity exceeded")
        out = dict(H=H.value, E=E.value, EA=EA, F=F.reshape(N, 3), U=U, X=Xw.reshape(N, 3), nn=nn)
        if want_g:
            out["G"] = G.reshape(N, self.D)
            out["dEdG"] = g.reshape(N, self.D)
        return out

    def mlp(self, sp, G):
        G = np.ascontiguousarray(G, dtype=np.float64)
        g = np.zeros_like(G)
        e = lib().orc_mlp(self._h, int(sp), _dp(G), _dp(g))
        return e, g


def neighbors(lat, x, rc, nd=3, xt=1, cap=4096):
    """Reference-order neighbour list: (counts[N], entries[M,4] of (j, ic0, ic1, ic2), wrapped X)."""
    lat = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    N = x.size // 3
    counts = np.zeros(N, dtype=np.int32)
    Xw = np.zeros(3 * N)
    total_cap = N * cap
    ent = np.zeros(4 * min(total_cap, 64 * 1024 * 1024), dtype=np.int32)
    rcode = lib().orc_neighbors(N, nd, xt, _dp(lat), _dp(x), rc, cap, _ip(counts), _ip(ent), ent.size // 4, _dp(Xw))
    if rcode != 0:
        raise RuntimeError("oracle: neighbour capacity exceeded")
    M = int(counts.sum())
    return counts, ent[:4 * M].reshape(M, 4), Xw.reshape(N, 3)


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ------------------------------------------------------------------ reference build
def have_reference() -> bool:
    return os.path.exists(REF_EXE)


def run_reference(model_path, spc, lat, x, nd=3, xt=1, nm=500, urep=0.0, p_gpa=0.0, repeats=1, threads=1,
                  workdir=None, timeout=3600):
    """Run the reference's CALL_MAISE (compiled from its sources) on one structure; parse its dump."""
    spc = np.asarray(spc, dtype=np.int64)
    lat = np.asarray(lat, dtype=np.float64).reshape(9)
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    N = spc.size
    tmp = workdir or tempfile.mkdtemp(prefix="maise_ref_")
    mp = os.path.join(tmp, "model")
    if os.path.abspath(model_path) != os.path.abspath(mp):
        with open(model_path) as src, open(mp, "w") as dst:
            dst.write(src.read())
    with open(os.path.join(tmp, "input"), "w") as f:
        f.write(f"{N} {nd} {xt} {nm} {float(urep)!r} {float(p_gpa)!r}\n")
        f.write(" ".join(repr(float(v)) for v in lat) + "\n")
        for i in range(N):
            f.write(f"{int(spc[i])} {float(x[i, 0])!r} {float(x[i, 1])!r} {float(x[i, 2])!r}\n")
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    res = subprocess.run([REF_EXE, "input", "output", str(repeats), str(threads)], cwd=tmp, env=env,
                         capture_output=True, text=True, timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError(f"reference run failed ({res.returncode}): {res.stderr[-2000:]}")
    out = {}
    for line in res.stdout.splitlines():
        if line.startswith("TIME"):
            out["time"] = float(line.split()[1])
    with open(os.path.join(tmp, "output")) as f:
        lines = f.read().split("\n")
    it = iter(lines)
    for line in it:
        if not line:
            continue
        tag = line.split()[0]
        if tag in ("H", "E"):
            out[tag] = float(line.split()[1])
        elif tag == "EA":
            out["EA"] = np.array([float(next(it)) for _ in range(N)])
        elif tag == "F":
            out["F"] = np.array([[float(v) for v in next(it).split()] for _ in range(N)])
        elif tag == "U":
            out["U"] = np.array([float(next(it)) for _ in range(6)])
        elif tag == "NN":
            out["nn"] = np.array([int(next(it)) for _ in range(N)])
        elif tag == "NBR":
            M = int(out["nn"].sum())
            out["nbr"] = np.array([[int(v) for v in next(it).split()] for _ in range(M)], dtype=np.int64).reshape(M, 4)
        elif tag == "G":
            out["G"] = np.array([[float(v) for v in next(it).split()] for _ in range(N)])
        elif tag == "X":
            out["X"] = np.array([[float(v) for v in next(it).split()] for _ in range(N)])
    return out


def fire_relax(model: "OracleModel", spc, lat, x, fmax=0.01, max_steps=2000, dt0=0.1, dtmax=1.0, maxstep=0.2,
               urep=0.0):
    """FIRE (Bitzek et al., PRL 97, 170201 (2006)) with unit masses, per-structure step cap, on FP64 oracle
    forces: the restatement the GPU batched relaxation is checked against.  Returns (x, E, steps, converged)."""
    x = np.array(x, dtype=np.float64).reshape(-1, 3).copy()
    v = np.zeros_like(x)
    dt, alpha, npos = dt0, 0.1, 0
    for it in range(max_steps):
        o = model.evaluate(spc, lat, x, urep=urep, threshold=False)
        F = o["F"]
        x = o["X"]
        if np.max(np.sum(F * F, axis=1)) < fmax * fmax:
            return x, o["E"], it, True
        vf = float(np.sum(F * v))
        if vf > 0:
            ff, vv = float(np.sum(F * F)), float(np.sum(v * v))
            v = (1 - alpha) * v + (alpha * np.sqrt(vv / ff) if ff > 0 else 0.0) * F
            if npos > 5:
                dt = min(dt * 1.1, dtmax)
                alpha *= 0.99
            npos += 1
        else:
            v[:] = 0.0
            alpha = 0.1
            dt *= 0.5
            npos = 0
        v = v + dt * F
        dr = dt * v
        nrm = np.sqrt(np.sum(dr * dr))
        if nrm > maxstep:
            dr *= maxstep / nrm
        x = x + dr
    o = model.evaluate(spc, lat, x, urep=urep)
    return o["X"], o["E"], max_steps, False


def lbfgs_direction(S, Y, gamma, q):
    """H q by the L-BFGS two-loop recursion (Nocedal & Wright, Numerical Optimization, Alg. 7.4): pairs (s, y) oldest
    first, rho = 1 / (y.s), initial inverse Hessian gamma I."""
    q = np.array(q, dtype=np.float64).copy()
    a = np.zeros(len(S))
    for i in range(len(S) - 1, -1, -1):
        a[i] = (S[i] @ q) / (Y[i] @ S[i])
        q -= a[i] * Y[i]
    z = gamma * q
    for i in range(len(S)):
        b = (Y[i] @ z) / (Y[i] @ S[i])
        z += S[i] * (a[i] - b)
    return z


def lbfgs_relax(model: "OracleModel", spc, lat, x, fmax=0.01, max_steps=2000, maxstep=0.2, memory=20, h0=1.0 / 70.0,
                urep=0.0, return_x_after=None):
    """Limited-memory BFGS without a line search on FP64 oracle forces, the restatement the GPU batched L-BFGS
    (mg_relax_lbfgs) is checked against: the newest (s, y) = (last step, F_old - F_new) pair is kept when
    y.s > 1e-10 |y||s| (at most `memory` pairs), gamma = s.y / y.y of the newest kept pair (h0 before any), direction
    p = H F; p.F <= 0 drops the memory and takes p = h0 F; the step is scaled so no atom moves more than maxstep.
    Returns (x, E, steps, converged); with return_x_after = k, the (unwrapped) positions after k steps instead."""
    x = np.array(x, dtype=np.float64).reshape(-1, 3).copy()
    S, Y = [], []
    gamma = h0
    f0 = dr = None
    for it in range(max_steps):
        if return_x_after is not None and it == return_x_after:
            return x
        o = model.evaluate(spc, lat, x, urep=urep, threshold=False)
        F = o["F"].reshape(-1)
        if np.max(np.sum(o["F"] * o["F"], axis=1)) < fmax * fmax:
            return o["X"], o["E"], it, True
        if f0 is not None:
            y = f0 - F
            ys, yy, ss = float(y @ dr), float(y @ y), float(dr @ dr)
            if yy > 0.0 and ys > 1e-10 * np.sqrt(yy * ss):
                S.append(dr.copy())
                Y.append(y)
                if len(S) > memory:
                    S.pop(0)
                    Y.pop(0)
                gamma = ys / yy
        p = lbfgs_direction(S, Y, gamma if S else h0, F)
        if not float(p @ F) > 0.0:
            S, Y = [], []
            gamma = h0
            p = h0 * F
        l2 = float(np.max(np.sum(p.reshape(-1, 3) ** 2, axis=1)))
        if l2 > maxstep * maxstep:
            p = p * (maxstep / np.sqrt(l2))
        x = x + p.reshape(-1, 3)
        dr, f0 = p, F
    if return_x_after is not None:
        return x
    o = model.evaluate(spc, lat, x, urep=urep)
    return o["X"], o["E"], max_steps, False
