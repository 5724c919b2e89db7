"""maise_amd -- MI355X-native evaluation of MAISE neural-network potentials.

Python mirror of the C ABI in include/maise_gpu.h (libmaise_gpu.so, loaded with ctypes).  The library is the
product: every call here goes to the HIP kernels; there is no CPU fallback, and importing the bindings fails
loudly when the library has not been built (python -m maise_amd.build, or __graft_entry__.build()).

    m = Model("models/nn_Cu_d3_v1")                 # READ_MODEL + READ_ANN + Build_PRS
    c = Context(m)                                  # one device stream
    c.set_structure(spc, lat, x, nd=3, xt=1)         # CAST_CELL
    r = c.evaluate()                                # CELL_FRC: {"E", "EA", "F", "U", "X"}
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmaise_gpu.so")

MG_EVAL_FORCES = 1
MG_EVAL_KEEP_G = 2
MG_PREC_FP32 = 0
MG_PREC_BF16 = 1
MG_PREC_BF16X3 = 2
EV2GPA = 160.2176565

_lib = None


class MaiseError(RuntimeError):
    pass


def lib():
    """Load libmaise_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MaiseError(f"{LIB_PATH} is missing: build it with `python -m maise_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp, fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float)
        sig = {
            "mg_last_error": ([], C.c_char_p),
            "mg_model_load": ([C.c_char_p, C.POINTER(vp)], C.c_int),
            "mg_model_free": ([vp], None),
            "mg_model_info": ([vp, ip, dp], C.c_int),
            "mg_model_id": ([vp, C.c_char_p, C.c_int], C.c_int),
            "mg_ctx_create": ([vp, C.c_int, C.POINTER(vp)], C.c_int),
            "mg_ctx_free": ([vp], None),
            "mg_ctx_set_urep": ([vp, C.c_double], C.c_int),
            "mg_ctx_set_precision": ([vp, C.c_int], C.c_int),
            "mg_ctx_set_timing": ([vp, C.c_int], C.c_int),
            "mg_ctx_set_wpb": ([vp, C.c_int, C.c_int], C.c_int),
            "mg_ctx_autotune": ([vp, C.c_int, ip], C.c_int),
            "mg_ctx_autotune_shape": ([vp, C.c_int, ip], C.c_int),
            "mg_ctx_get_shape": ([vp, ip], C.c_int),
            "mg_ctx_set_wpb_nbr": ([vp, C.c_int], C.c_int),
            "mg_ctx_set_autotune": ([vp, C.c_int], C.c_int),
            "mg_relax_log": ([vp, C.c_int, dp], C.c_int),
            "mg_ctx_set_skin": ([vp, C.c_double], C.c_int),
            "mg_ctx_skin_builds": ([vp, ip], C.c_int),
            "mg_ctx_set_graph": ([vp, C.c_int, C.c_double], C.c_int),
            "mg_ctx_set_pair_store": ([vp, C.c_int], C.c_int),
            "mg_set_structure": ([vp, C.c_int, C.c_int, C.c_int, ip, dp, dp], C.c_int),
            "mg_set_batch": ([vp, C.c_int, ip, ip, dp, ip, dp], C.c_int),
            "mg_eval": ([vp, C.c_int], C.c_int),
            "mg_get": ([vp, C.c_int, dp, dp, dp, dp, dp], C.c_int),
            "mg_get_batch": ([vp, dp, dp, dp], C.c_int),
            "mg_get_batch_x": ([vp, dp], C.c_int),
            "mg_calibrate_rate": ([C.c_int, C.c_int, dp], C.c_int),
            "mg_eval_static": ([vp, C.c_int, C.c_int, C.c_int, ip, dp, dp, C.c_double, dp, dp, dp, dp], C.c_int),
            "mg_get_neighbors": ([vp, ip, ip, ip, ip], C.c_int),
            "mg_get_descriptors": ([vp, fp], C.c_int),
            "mg_get_timing": ([vp, fp], C.c_int),
            "mg_md_init": ([vp, C.c_double, C.c_ulonglong], C.c_int),
            "mg_md_set_velocities": ([vp, dp], C.c_int),
            "mg_md_get_velocities": ([vp, dp], C.c_int),
            "mg_md_run": ([vp, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp], C.c_int),
            "mg_md_run_npt": ([vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp],
                              C.c_int),
            "mg_relax": ([vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, ip, ip], C.c_int),
            "mg_relax_lbfgs": ([vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, ip, ip], C.c_int),
            "mg_relax_cell": ([vp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                               ip, ip], C.c_int),
            "mg_get_lattice": ([vp, C.c_int, dp], C.c_int),
            "mg_sur_nparams": ([C.c_int], C.c_int),
            "mg_sur_create": ([C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)], C.c_int),
            "mg_sur_free": ([vp], None),
            "mg_sur_get_params": ([vp, fp], C.c_int),
            "mg_sur_set_params": ([vp, fp], C.c_int),
            "mg_sur_set_norm": ([vp, fp, fp, C.c_float, C.c_float], C.c_int),
            "mg_sur_train": ([vp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong, fp], C.c_int),
            "mg_sur_train_begin": ([vp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_ulonglong], C.c_int),
            "mg_sur_train_end": ([vp, fp], C.c_int),
            "mg_sur_grad": ([vp, fp, fp, C.c_int, fp, fp], C.c_int),
            "mg_sur_predict": ([vp, fp, C.c_int, fp], C.c_int),
            "mg_features": ([vp, fp], C.c_int),
            "mg_evo_offspring": ([C.c_int, C.c_int, C.c_int, ip, dp, dp, dp, C.c_int, C.c_double, C.c_double,
                                  C.c_ulonglong, C.c_int, ip, dp, dp, ip], C.c_int),
            "mg_evo_min_distance": ([dp, dp, C.c_int], C.c_double),
            "mg_evo_mutate": ([C.c_int, C.c_int, C.c_int, ip, dp, dp, dp, C.c_int, C.c_double, C.c_double, C.c_double,
                               C.c_double, C.c_ulonglong, C.c_int, ip, dp, dp, ip], C.c_int),
            "mg_evo_fingerprint": ([C.c_int, ip, C.c_int, ip, dp, dp, C.c_int, C.c_double, C.c_double, C.c_double,
                                    fp], C.c_int),
            "mg_evo_unique": ([C.c_int, dp, dp, fp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, ip],
                              C.c_int),
            "mg_poscar_read": ([C.c_char_p, C.c_int, C.c_int, C.c_char_p, dp, ip, ip, ip, ip, dp, ip, dp, ip],
                               C.c_int),
            "mg_poscar_write": ([C.c_char_p, C.c_char_p, C.c_double, dp, C.c_int, C.c_int, ip, ip, C.c_int, dp, ip,
                                 C.c_int, dp], C.c_int),
            "mg_outcar_write": ([C.c_char_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_int, ip, ip, C.c_int, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_int],
                                C.c_int),
            "mg_element_symbol": ([C.c_int], C.c_char_p),
            "mg_element_z": ([C.c_char_p], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().mg_last_error()
        raise MaiseError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class Model:
    """An nn_* model loaded by the GPU library (format of src/nutl.c:394-498)."""

    def __init__(self, path: str):
        self._h = C.c_void_p()
        _check(lib().mg_model_load(os.fsencode(path), C.byref(self._h)), f"loading model {path}")
        info = np.zeros(16, dtype=np.int32)
        rc = C.c_double()
        _check(lib().mg_model_info(self._h, _ip(info), C.byref(rc)), "mg_model_info")
        self.nspc, self.D, self.H1, self.H2, self.nsym, self.ng2, self.ng4, self.nw = (int(v) for v in info[:8])
        self.spcz = [int(v) for v in info[8:8 + self.nspc]]
        self.g4_variant, self.shared_eta, self.mfma = int(info[12]), bool(info[13]), bool(info[14])
        self.rc = rc.value
        self.path = path
        mid = C.create_string_buffer(64)
        _check(lib().mg_model_id(self._h, mid, 64), "mg_model_id")
        self.id = mid.value.decode()

    def __del__(self):
        try:
            if self._h:
                lib().mg_model_free(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def calibrate_rate(op: int, device: int = 0) -> float:
    """Measured issue rate (lane operations/s) of op: 0 FP32 FMA, 1 exp2, 2 cos, 3 sqrt, 4 rcp."""
    v = C.c_double()
    _check(lib().mg_calibrate_rate(device, op, C.byref(v)), "mg_calibrate_rate")
    return v.value


def eval_static(model: "Model", spc, lat, x, nd: int = 3, xt: int = 1, urep: float = 0.0) -> dict:
    """FP32 E/F/virial of one structure through mg_eval_static, the entry point that allocates nothing (all
    intermediate buffers in the library's fixed per-device scratch array)."""
    spc = np.ascontiguousarray(spc, dtype=np.int32)
    lat = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    N = spc.size
    E = np.zeros(1)
    EA = np.zeros(N)
    F = np.zeros((N, 3))
    U = np.zeros(6)
    _check(lib().mg_eval_static(model._h, N, nd, xt, _ip(spc), _dp(lat), _dp(x), float(urep), _dp(E), _dp(EA), _dp(F),
                                _dp(U)), "mg_eval_static")
    return {"E": float(E[0]), "EA": EA, "F": F, "U": U}


class Context:
    """One evaluation context: device buffers and one HIP stream; a single structure or a batch."""

    def __init__(self, model: Model, device: int = 0, urep: float = 0.0, precision: int = MG_PREC_FP32):
        self.model = model
        self._h = C.c_void_p()
        _check(lib().mg_ctx_create(model._h, device, C.byref(self._h)), "mg_ctx_create")
        self.set_urep(urep)
        _check(lib().mg_ctx_set_precision(self._h, precision), "mg_ctx_set_precision")
        self.N = 0
        self.S = 0
        self._atom_off = None

    def close(self):
        if self._h:
            lib().mg_ctx_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_urep(self, urep: float):
        _check(lib().mg_ctx_set_urep(self._h, float(urep)), "mg_ctx_set_urep")

    def set_precision(self, prec: int):
        _check(lib().mg_ctx_set_precision(self._h, int(prec)), "mg_ctx_set_precision")

    def set_timing(self, on: bool):
        _check(lib().mg_ctx_set_timing(self._h, 1 if on else 0), "mg_ctx_set_timing")

    def set_wpb(self, wpb_desc: int = 0, wpb_force: int = 0):
        """Atoms (waves) per workgroup of the descriptor and force kernels (0 = largest that fits 64 KiB LDS)."""
        _check(lib().mg_ctx_set_wpb(self._h, int(wpb_desc), int(wpb_force)), "mg_ctx_set_wpb")

    def set_graph(self, enable: bool = True, headroom: float = 0.125):
        """Relaxation/MD replay a captured HIP graph per iteration (edge buffers sized with this headroom;
        overflow is detected and rolled back, so results do not depend on it); enable=False synchronises every
        iteration."""
        _check(lib().mg_ctx_set_graph(self._h, 1 if enable else 0, float(headroom)), "mg_ctx_set_graph")

    def set_pair_store(self, enable: bool = True):
        """The descriptor kernel stores its G4 neighbour-pair lists for the force kernel (default); False makes the
        force kernel enumerate the pairs itself.  Results do not depend on it."""
        _check(lib().mg_ctx_set_pair_store(self._h, 1 if enable else 0), "mg_ctx_set_pair_store")

    def autotune(self, reps: int = 3):
        """Tune the launch shape on the current structure or batch (neighbour fill, descriptor and force waves per
        workgroup, stored pair lists or enumeration); returns (nbr, desc, force, pair_store)."""
        out = np.zeros(4, dtype=np.int32)
        _check(lib().mg_ctx_autotune_shape(self._h, int(reps), _ip(out)), "mg_ctx_autotune_shape")
        return tuple(int(v) for v in out)

    def shape(self):
        """Current launch shape (nbr, desc, force waves per workgroup, pair_store); 0 = automatic."""
        out = np.zeros(4, dtype=np.int32)
        _check(lib().mg_ctx_get_shape(self._h, _ip(out)), "mg_ctx_get_shape")
        return tuple(int(v) for v in out)

    def set_wpb_nbr(self, wpb: int = 0):
        _check(lib().mg_ctx_set_wpb_nbr(self._h, int(wpb)), "mg_ctx_set_wpb_nbr")

    def set_skin(self, skin: float = 1.0):
        """Skin neighbour lists (A; 0 = off) for fixed-cell relaxation and MD: candidates within rc + skin are kept
        until an atom has moved skin/2, and every step applies the reference's FP64 neighbour test to them, so the
        results are bit-identical to rebuilding the lists every step."""
        _check(lib().mg_ctx_set_skin(self._h, float(skin)), "mg_ctx_set_skin")

    def skin_builds(self) -> int:
        """Skin-candidate builds over the context's life."""
        n = np.zeros(1, dtype=np.int32)
        _check(lib().mg_ctx_skin_builds(self._h, _ip(n)), "mg_ctx_skin_builds")
        return int(n[0])

    def set_autotune(self, on: bool = True):
        """Relaxation and MD tune the launch shape before their first captured iteration (results unchanged)."""
        _check(lib().mg_ctx_set_autotune(self._h, 1 if on else 0), "mg_ctx_set_autotune")

    def set_structure(self, spc, lat, x, nd: int = 3, xt: int = 1):
        spc = np.ascontiguousarray(spc, dtype=np.int32)
        lat = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != 3 * spc.size:
            raise ValueError("x must hold 3 coordinates per atom")
        _check(lib().mg_set_structure(self._h, spc.size, nd, xt, _ip(spc), _dp(lat), _dp(x)), "mg_set_structure")
        self.N, self.S = spc.size, 1
        self._atom_off = np.array([0, spc.size], dtype=np.int32)

    def set_batch(self, structures):
        """structures: list of (spc, lat, x_cartesian[, nd]) tuples."""
        offs, nds, lats, spcs, xs = [0], [], [], [], []
        for s in structures:
            spc, lat, x = s[0], s[1], s[2]
            nd = s[3] if len(s) > 3 else 3
            spcs.append(np.asarray(spc, dtype=np.int32))
            lats.append(np.asarray(lat, dtype=np.float64).reshape(9))
            xs.append(np.asarray(x, dtype=np.float64).reshape(-1))
            nds.append(nd)
            offs.append(offs[-1] + len(spc))
        off = np.array(offs, dtype=np.int32)
        nd = np.array(nds, dtype=np.int32)
        lat = np.ascontiguousarray(np.concatenate(lats))
        spc = np.ascontiguousarray(np.concatenate(spcs))
        x = np.ascontiguousarray(np.concatenate(xs))
        _check(lib().mg_set_batch(self._h, len(structures), _ip(off), _ip(nd), _dp(lat), _ip(spc), _dp(x)), "mg_set_batch")
        self.N, self.S = int(off[-1]), len(structures)
        self._atom_off = off

    def eval(self, forces: bool = True, keep_g: bool = False):
        flags = (MG_EVAL_FORCES if forces else 0) | (MG_EVAL_KEEP_G if keep_g else 0)
        _check(lib().mg_eval(self._h, flags), "mg_eval")

    def result(self, s: int = 0, forces: bool = True) -> dict:
        n = int(self._atom_off[s + 1] - self._atom_off[s])
        E = C.c_double()
        EA = np.zeros(n)
        X = np.zeros(3 * n)
        F = np.zeros(3 * n) if forces else None
        U = np.zeros(6) if forces else None
        _check(lib().mg_get(self._h, s, C.byref(E), _dp(EA), _dp(F), _dp(U), _dp(X)), "mg_get")
        out = {"E": E.value, "EA": EA, "X": X.reshape(n, 3)}
        if forces:
            out["F"] = F.reshape(n, 3)
            out["U"] = U
        return out

    def evaluate(self, forces: bool = True) -> dict:
        self.eval(forces=forces)
        return self.result(0, forces=forces)

    def batch_result(self, forces: bool = True) -> dict:
        E = np.zeros(self.S)
        U = np.zeros(6 * self.S) if forces else None
        F = np.zeros(3 * self.N) if forces else None
        _check(lib().mg_get_batch(self._h, _dp(E), _dp(U), _dp(F)), "mg_get_batch")
        out = {"E": E}
        if forces:
            out["U"] = U.reshape(self.S, 6)
            out["F"] = F.reshape(self.N, 3)
        return out

    def batch_positions(self) -> np.ndarray:
        """Wrapped positions [N, 3] of every structure of the batch, in batch order (one device copy)."""
        X = np.zeros(3 * self.N)
        _check(lib().mg_get_batch_x(self._h, _dp(X)), "mg_get_batch_x")
        return X.reshape(self.N, 3)

    def neighbors(self):
        """(row_off[N+1], j[M], img[M,3]) of the last evaluation; rows sorted by (species, j, image)."""
        M = C.c_int()
        _check(lib().mg_get_neighbors(self._h, C.byref(M), None, None, None), "mg_get_neighbors")
        row = np.zeros(self.N + 1, dtype=np.int32)
        j = np.zeros(max(M.value, 1), dtype=np.int32)
        img = np.zeros(3 * max(M.value, 1), dtype=np.int32)
        _check(lib().mg_get_neighbors(self._h, C.byref(M), _ip(row), _ip(j), _ip(img)), "mg_get_neighbors")
        return row, j[:M.value], img[:3 * M.value].reshape(-1, 3)

    def descriptors(self) -> np.ndarray:
        G = np.zeros(self.N * self.model.D, dtype=np.float32)
        _check(lib().mg_get_descriptors(self._h, G.ctypes.data_as(C.POINTER(C.c_float))), "mg_get_descriptors")
        return G.reshape(self.N, self.model.D)

    def timing(self) -> np.ndarray:
        """Stage times (ms), averaged over the evaluations since timing was enabled or last read."""
        t = np.zeros(6, dtype=np.float32)
        _check(lib().mg_get_timing(self._h, t.ctypes.data_as(C.POINTER(C.c_float))), "mg_get_timing")
        return t


    # ------------------------------------------------------------- dynamics
    def md_init(self, T: float, seed: int = 1):
        """Seeded Maxwell-Boltzmann velocities at T (K), zero centre-of-mass momentum, exact T."""
        _check(lib().mg_md_init(self._h, float(T), int(seed)), "mg_md_init")

    def md_set_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        _check(lib().mg_md_set_velocities(self._h, _dp(v)), "mg_md_set_velocities")

    def md_velocities(self) -> np.ndarray:
        v = np.zeros(3 * self.N)
        _check(lib().mg_md_get_velocities(self._h, _dp(v)), "mg_md_get_velocities")
        return v.reshape(self.N, 3)

    def md_run(self, nsteps: int, dt: float = 1.0, T: float = 300.0, tau: float = 100.0):
        """NVT (tau > 0) or NVE (tau <= 0) velocity Verlet on the device; returns per-step (Epot, Ekin) in eV."""
        ep, ek = np.zeros(max(nsteps, 1)), np.zeros(max(nsteps, 1))
        _check(lib().mg_md_run(self._h, int(nsteps), float(dt), float(T), float(tau), _dp(ep), _dp(ek)), "mg_md_run")
        return ep[:nsteps], ek[:nsteps]

    def md_run_npt(self, nsteps: int, dt: float = 1.0, T: float = 300.0, tau: float = 100.0, p_gpa: float = 0.0,
                   ktt: float = 0.005 / 50.0):
        """NPT (tau > 0) or isobaric NVE (tau <= 0) with the Berendsen barostat (ktt = compressibility / coupling
        time, 1/(GPa fs)); returns per-step (Epot eV, Ekin eV, volume A^3, internal pressure GPa)."""
        n = max(nsteps, 1)
        ep, ek, vol, pi = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
        _check(lib().mg_md_run_npt(self._h, int(nsteps), float(dt), float(T), float(tau), float(p_gpa), float(ktt),
                                   _dp(ep), _dp(ek), _dp(vol), _dp(pi)), "mg_md_run_npt")
        return ep[:nsteps], ek[:nsteps], vol[:nsteps], pi[:nsteps]

    def features(self) -> np.ndarray:
        """Per-structure surrogate inputs of the last evaluation: [S, nspc * D], per species the mean of the
        descriptor rows of that species' atoms."""
        m = self.model
        f = np.zeros((self.S, m.nspc * m.D), dtype=np.float32)
        _check(lib().mg_features(self._h, _fp(f)), "mg_features")
        return f

    def relax(self, max_steps: int = 1000, fmax: float = 0.01, dt0: float = 0.1, dtmax: float = 1.0,
              maxstep: float = 0.2):
        """Batched FIRE over all structures of the context; returns (steps[S], converged[S])."""
        steps = np.zeros(self.S, dtype=np.int32)
        done = np.zeros(self.S, dtype=np.int32)
        _check(lib().mg_relax(self._h, int(max_steps), float(fmax), float(dt0), float(dtmax), float(maxstep),
                              _ip(steps), _ip(done)), "mg_relax")
        return steps, done.astype(bool)

    def relax_lbfgs(self, max_steps: int = 1000, fmax: float = 0.01, maxstep: float = 0.2, memory: int = 20,
                    h0: float = 1.0 / 70.0):
        """Batched L-BFGS over all structures of the context (memory pairs per structure, initial inverse Hessian h0
        in A^2/eV, no atom moving more than maxstep A per step); returns (steps[S], converged[S])."""
        steps = np.zeros(self.S, dtype=np.int32)
        done = np.zeros(self.S, dtype=np.int32)
        _check(lib().mg_relax_lbfgs(self._h, int(max_steps), float(fmax), float(maxstep), int(memory), float(h0),
                                    _ip(steps), _ip(done)), "mg_relax_lbfgs")
        return steps, done.astype(bool)

    def relax_log(self) -> np.ndarray:
        """The last relaxation's wall time by segment between host checks: rows of (iterations done, active
        structures, active atoms, seconds)."""
        n = lib().mg_relax_log(self._h, 0, None)
        if n < 0:
            _check(n, "mg_relax_log")
        rec = np.zeros(4 * max(n, 1))
        _check(min(0, lib().mg_relax_log(self._h, n, _dp(rec))), "mg_relax_log")
        return rec[:4 * n].reshape(n, 4)

    def relax_cell(self, p_gpa: float = 0.0, max_steps: int = 3000, fmax: float = 0.01, smax_gpa: float = 0.01,
                   dt0: float = 0.1, dtmax: float = 1.0, maxstep: float = 0.2):
        """Batched FIRE over positions and lattice at pressure p (GPa) to every |F| < fmax (eV/A) and every stress
        component < smax (GPa); returns (steps[S], converged[S]); lattices via lattice(s)."""
        steps = np.zeros(self.S, dtype=np.int32)
        done = np.zeros(self.S, dtype=np.int32)
        _check(lib().mg_relax_cell(self._h, float(p_gpa), int(max_steps), float(fmax), float(smax_gpa), float(dt0),
                                   float(dtmax), float(maxstep), _ip(steps), _ip(done)), "mg_relax_cell")
        return steps, done.astype(bool)

    def lattice(self, s: int = 0) -> np.ndarray:
        lat = np.zeros(9)
        _check(lib().mg_get_lattice(self._h, int(s), _dp(lat)), "mg_get_lattice")
        return lat.reshape(3, 3)


def call_maise_symbol():
    """The exported drop-in entry point (for ABI tests)."""
    return lib().CALL_MAISE

