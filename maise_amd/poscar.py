"""POSCAR/CONTCAR reading and writing (VASP-5 layout, as READ_CELL/SAVE_CELL use: src/util.c:820-1071).

Layout: comment; scale; three lattice rows; optional species symbols; counts; optional "Selective";
"Direct"/"Cartesian"; one position row per atom (optional T/F flags); optional velocities.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np


@dataclass
class Poscar:
    comment: str
    lattice: np.ndarray        # [3,3] rows, A (scale applied)
    symbols: List[str]
    counts: List[int]
    frac: np.ndarray           # [N,3] fractional coordinates
    velocities: Optional[np.ndarray] = None

    @property
    def species(self) -> np.ndarray:
        return np.repeat(np.arange(len(self.counts)), self.counts).astype(np.int32)

    @property
    def cart(self) -> np.ndarray:
        return self.frac @ self.lattice


def read_poscar(path: str) -> Poscar:
    with open(path) as f:
        lines = f.read().splitlines()
    comment = lines[0].strip()
    scale = float(lines[1].split()[0])
    lat = np.array([[float(v) for v in lines[2 + k].split()[:3]] for k in range(3)]) * scale
    li = 5
    toks = lines[li].split()
    symbols: List[str] = []
    if not toks[0][0].isdigit():
        symbols = toks
        li += 1
        toks = lines[li].split()
    counts = [int(v) for v in toks]
    li += 1
    if lines[li].strip()[:1] in ("S", "s"):
        li += 1
    cartesian = lines[li].strip()[:1] in ("C", "c", "K", "k")
    li += 1
    N = sum(counts)
    pos = np.array([[float(v) for v in lines[li + i].split()[:3]] for i in range(N)])
    if cartesian:
        pos = (pos * scale) @ np.linalg.inv(lat)
    vel = None
    rest = [l for l in lines[li + N:] if l.strip()]
    if len(rest) >= N:
        try:
            vel = np.array([[float(v) for v in rest[i].split()[:3]] for i in range(N)])
        except ValueError:
            vel = None
    return Poscar(comment, lat, symbols, counts, pos, vel)


def write_poscar(path: str, p: Poscar) -> None:
    out = [p.comment, "   1.00000000000000"]
    for r in p.lattice:
        out.append("  " + "".join(f"{v:22.16f}" for v in r))
    if p.symbols:
        out.append("   " + "   ".join(p.symbols))
    out.append("   " + "   ".join(str(c) for c in p.counts))
    out.append("Direct")
    for r in p.frac:
        out.append("".join(f"{v:20.16f}" for v in r))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
