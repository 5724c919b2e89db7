"""POSCAR/CONTCAR and OUTCAR/OSZICAR files through the library's readers and writers (mg_poscar_read,
mg_poscar_write, mg_outcar_write in include/maise_gpu.h), which follow READ_CELL / SAVE_CELL (src/util.c:820-1030)
and CELL_OUT (src/cmod.c:618-687).

Layout read: comment; scale; three lattice rows; species symbols (required, as the reference requires them); counts;
optional "Selective dynamics"; "Direct" or Cartesian (taken as written, unscaled, as READ_CELL does); one position
row per atom with optional T/F flags; optional velocities.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _check, _dp, _ip, lib


@dataclass
class Poscar:
    comment: str
    lattice: np.ndarray        # [3,3] rows, A (scale applied)
    symbols: List[str]
    counts: List[int]
    x: np.ndarray              # [N,3] Cartesian, A
    ff: np.ndarray             # [N,3] 1 = free, 0 = fixed (selective dynamics)
    velocities: Optional[np.ndarray] = None

    @property
    def species(self) -> np.ndarray:
        """Species index per atom (position of its symbol in the file's species line)."""
        return np.repeat(np.arange(len(self.counts)), self.counts).astype(np.int32)

    @property
    def z(self) -> np.ndarray:
        return np.array([element_z(s) for s in self.symbols], dtype=np.int32)

    @property
    def cart(self) -> np.ndarray:
        return self.x

    @property
    def frac(self) -> np.ndarray:
        return self.x @ np.linalg.inv(self.lattice)


def element_symbol(z: int) -> str:
    return lib().mg_element_symbol(int(z)).decode()


def element_z(symbol: str) -> int:
    return int(lib().mg_element_z(symbol.encode()))


def read_poscar(path: str, nd: int = 3) -> Poscar:
    """Read a POSCAR/CONTCAR; nd = 0 (cluster) centres the atoms on the origin as the reference does."""
    title = C.create_string_buffer(256)
    lat = np.zeros(9)
    nspc, natoms = C.c_int(0), C.c_int(0)
    spcz = np.zeros(10, dtype=np.int32)
    spcn = np.zeros(10, dtype=np.int32)
    p = path.encode()
    _check(lib().mg_poscar_read(p, int(nd), 0, title, _dp(lat), C.byref(nspc), _ip(spcz), _ip(spcn), C.byref(natoms),
                                None, None, None, None), "mg_poscar_read")
    N = natoms.value
    x = np.zeros((max(N, 1), 3))
    ff = np.ones((max(N, 1), 3), dtype=np.int32)
    v = np.zeros((max(N, 1), 3))
    has_v = C.c_int(0)
    _check(lib().mg_poscar_read(p, int(nd), N, title, _dp(lat), C.byref(nspc), _ip(spcz), _ip(spcn),
                                C.byref(natoms), _dp(x), _ip(ff), _dp(v), C.byref(has_v)), "mg_poscar_read")
    ns = nspc.value
    return Poscar(title.value.decode(), lat.reshape(3, 3), [element_symbol(z) for z in spcz[:ns]],
                  [int(n) for n in spcn[:ns]], x[:N].copy(), ff[:N].copy(), v[:N].copy() if has_v.value else None)


def write_poscar(path: str, lat, z, counts, x, comment: str = "", scale: float = 1.0, nd: int = 3, ff=None,
                 sd: bool = False, velocities=None) -> None:
    """Write a CONTCAR in the reference's format: Cartesian positions x[N,3] in species order, atomic numbers z and
    counts per species, direct coordinates on output (clusters centred in the box)."""
    lat = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
    z = np.ascontiguousarray(z, dtype=np.int32)
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    ffa = None if ff is None else np.ascontiguousarray(ff, dtype=np.int32).reshape(-1, 3)
    va = None if velocities is None else np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1, 3)
    _check(lib().mg_poscar_write(path.encode(), comment.encode(), float(scale), _dp(lat), int(nd), len(z), _ip(z),
                                 _ip(counts), len(x), _dp(x), _ip(ffa), int(bool(sd)), _dp(va)), "mg_poscar_write")


def write_outcar(directory: str, it: int, lat, z, counts, E: float, H: float, x=None, F=None, EA=None, U=None,
                 version: str = "", model_id: str = "", mint: int = 0, rlxt: int = 0, p_gpa: float = 0.0,
                 nd: int = 3, detail: bool = True) -> None:
    """Append iteration `it` to directory/OUTCAR and directory/OSZICAR (the run header too when it == 0).
    U is the virial (eV, xx yy zz xy yz zx as the evaluation returns it); detail adds positions, forces, atom
    energies and the stress block."""
    lat = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
    z = np.ascontiguousarray(z, dtype=np.int32)
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    arr = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (x, F, EA, U)]
    _check(lib().mg_outcar_write(directory.encode(), int(it), version.encode(), model_id.encode(), int(mint),
                                 int(rlxt), float(p_gpa), int(nd), len(z), _ip(z), _ip(counts), int(counts.sum()),
                                 _dp(lat), _dp(arr[0]), _dp(arr[1]), _dp(arr[2]), _dp(arr[3]), float(E), float(H),
                                 int(bool(detail))), "mg_outcar_write")
