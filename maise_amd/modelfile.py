"""Writer and reader helpers for MAISE ``nn_*`` model files.

The text layout follows the reference model writer ``SAVE_ANN`` (src/nutl.c:394-498) and the readers
``READ_MODEL``/``READ_ANN``/``Build_PRS`` (src/util.c:1233-1335, src/nutl.c:502-633, src/nprs.c:75-257):
a header of 25-character tagged lines, the per-species input ranges (min, max, range), the weight vector in
the ``V2W`` order (src/nmlp.c:591-608: per species, all biases layer by layer, then the weights of each layer
as [out][in]), and the basis block appended after ``  B2A          Scaling``.

The writer is used for synthetic models (seeded weights, project-chosen bases); any file it writes is read
unchanged by the reference and by this library.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

_ELEMENTS = (
    "X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am"
).split()


def symbol(z: int) -> str:
    return _ELEMENTS[z]


def atomic_number(sym: str) -> int:
    return _ELEMENTS.index(sym)


@dataclass
class Basis:
    """A Behler-Parrinello basis in the reference's table form.

    ``funcs`` rows are (type, Rc-idx, Rs-idx, n1-idx, n2-idx, k-idx, z-idx, l-idx); ``params`` maps the
    seven parameter names (Rc, Rs, n1, n2, k, z, l) to value lists.  ``b2a`` and ``scale`` multiply distances
    (Rc, Rs) and divide the inverse-square parameters (n1, n2) squared, as src/nprs.c:183-189 does.
    """

    funcs: List[Sequence[int]]
    params: dict
    b2a: float = 1.0
    scale: float = 1.0

    @property
    def ng2(self) -> int:
        return sum(1 for f in self.funcs if f[0] == 2)

    @property
    def ng4(self) -> int:
        return sum(1 for f in self.funcs if f[0] == 4)

    def width(self, nspc: int) -> int:
        """Descriptor width NG2*S + NG4*S(S+1)/2 (src/nprs.c:213-220)."""
        return self.ng2 * nspc + self.ng4 * nspc * (nspc + 1) // 2

    def text(self) -> str:
        lines = ["  B2A          Scaling", f"  {self.b2a:.9g}  {self.scale:.9g}", "",
                 "  N  GN  Rc  Rs   n1  n2  k   z   l", ""]
        for i, f in enumerate(self.funcs):
            lines.append(f"{i + 1:3d}   " + "   ".join(f"{v:d}" for v in f))
        lines.append("")
        for name in ("Rc", "Rs", "n1", "n2", "k", "z", "l"):
            vals = self.params[name]
            lines.append(f" {name:<4s}{len(vals):d}   " + "   ".join(f"{v:.10g}" for v in vals))
        return "\n".join(lines) + "\n"


def standard_basis(rc: float = 6.0, eta2=None, eta4=None, zetas=(1.0, 2.0, 4.0, 16.0),
                   lambdas=(-1.0, 1.0), b2a: float = 1.0, scale: float = 1.0) -> Basis:
    """8 G2 + 43 G4 basis with the selection pattern the shipped v0/v1 models use.

    G4 (zeta, lambda) combinations per eta index: eta 0-2 take zeta in {1st, 2nd} x both lambdas, eta 3-5
    take all eight, eta 6 takes seven (dropping the largest zeta with lambda=+1).  This is the 43-function
    table of models/nn_Cu_d3_v1.
    """
    if eta2 is None:
        eta2 = np.geomspace(0.003, 0.5, 8)
    if eta4 is None:
        eta4 = np.geomspace(0.003, 0.3, 7)
    funcs = [(2, 0, 0, n, 0, 0, 0, 0) for n in range(len(eta2))]
    for e in range(len(eta4)):
        if e < 3:
            combos = [(z, l) for z in range(2) for l in range(2)]
        elif e < 6:
            combos = [(z, l) for z in range(4) for l in range(2)]
        else:
            combos = [(z, l) for z in range(4) for l in range(2)][:7]
        for z, l in combos:
            funcs.append((4, 0, 0, 0, e, 0, z, l))
    params = {"Rc": [rc], "Rs": [0.0], "n1": list(map(float, eta2)), "n2": list(map(float, eta4)),
              "k": [1.0], "z": list(zetas), "l": list(lambdas)}
    return Basis(funcs=funcs, params=params, b2a=b2a, scale=scale)


def generic_basis(rc: float = 6.0, eta2=None, eta4=None, zetas=(1.0, 3.0), lambdas=(-1.0, 1.0), b2a: float = 1.0,
                  scale: float = 1.0) -> Basis:
    """A G2 + G4 basis on the full (eta, zeta, lambda) grid with zeta values of the caller's choosing (the
    default zeta = {1, 3} matches neither shipped G4 table)."""
    if eta2 is None:
        eta2 = np.geomspace(0.005, 0.4, 6)
    if eta4 is None:
        eta4 = np.geomspace(0.005, 0.2, 3)
    funcs = [(2, 0, 0, n, 0, 0, 0, 0) for n in range(len(eta2))]
    for e in range(len(eta4)):
        for z in range(len(zetas)):
            for l in range(len(lambdas)):
                funcs.append((4, 0, 0, 0, e, 0, z, l))
    params = {"Rc": [rc], "Rs": [0.0], "n1": list(map(float, eta2)), "n2": list(map(float, eta4)),
              "k": [1.0], "z": list(zetas), "l": list(lambdas)}
    return Basis(funcs=funcs, params=params, b2a=b2a, scale=scale)


@dataclass
class Network:
    """Per-species D-H1-H2-1 network: ``biases[s][k]`` (length NU[k+1]) and ``weights[s][k]`` ([out][in])."""

    spcz: List[int]
    arch: List[int]
    biases: list
    weights: list
    rmin: np.ndarray  # [S, D]
    rmax: np.ndarray  # [S, D]
    drange: np.ndarray  # [S, D] = 1/(max-min) scaling applied to the inputs
    info: dict = field(default_factory=dict)

    @property
    def nw(self) -> int:
        a = self.arch
        return len(self.spcz) * sum((a[k] + 1) * a[k + 1] for k in range(len(a) - 1))

    def vector(self) -> np.ndarray:
        out = []
        for s in range(len(self.spcz)):
            for k in range(len(self.arch) - 1):
                out.extend(np.asarray(self.biases[s][k], dtype=np.float64).ravel())
            for k in range(len(self.arch) - 1):
                out.extend(np.asarray(self.weights[s][k], dtype=np.float64).ravel())
        return np.asarray(out)


def random_network(spcz: Sequence[int], arch: Sequence[int], seed: int) -> Network:
    """Seeded Gaussian weights with std 1/sqrt(fan_in), zero biases; input ranges set later."""
    rng = np.random.default_rng(seed)
    biases, weights = [], []
    for _ in spcz:
        b, w = [], []
        for k in range(len(arch) - 1):
            b.append(np.zeros(arch[k + 1]))
            w.append(rng.normal(0.0, 1.0 / math.sqrt(arch[k]), size=(arch[k + 1], arch[k])))
        biases.append(b)
        weights.append(w)
    S, D = len(spcz), arch[0]
    return Network(spcz=list(spcz), arch=list(arch), biases=biases, weights=weights,
                   rmin=np.zeros((S, D)), rmax=np.ones((S, D)), drange=np.ones((S, D)))


def _row(tag: str, value: str) -> str:
    body = f"|  {tag:<21s}|  {value}"
    return body.ljust(76) + "|"


def write_model(path: str, net: Network, basis, model_id: str = "5EED0000",
                version: str = "maise.3.0.00") -> None:
    """Write an ``nn_*`` model file readable by the reference (format of src/nutl.c:394-498).

    ``basis`` is one Basis (used for every species) or a list of one Basis per species: species-specific blocks
    follow each other separated by one blank line, the layout src/nprs.c:154-231 reads (after a block it takes
    one line, then the next block's header line, then that block's scaling line)."""
    bar = "-" * 77
    S = len(net.spcz)
    D = net.arch[0]
    blocks = list(basis) if isinstance(basis, (list, tuple)) else [basis]
    if len(blocks) not in (1, S):
        raise ValueError("give one basis block, or one per species")
    for b in blocks:
        if b.width(S) != D or [tuple(f) for f in b.funcs] != [tuple(f) for f in blocks[0].funcs]:
            raise ValueError(f"basis width {b.width(S)} != input width {D}, or blocks with different functions")
    lines = [bar, "|                    neural network general information                     |", bar,
             _row("model unique ID", model_id),
             _row("number of species", str(S)),
             _row("species types", " ".join(f"{z:d}" for z in net.spcz)),
             _row("species names", " ".join(symbol(z) for z in net.spcz)),
             _row("number of layers", str(len(net.arch))),
             _row("architecture", " ".join(str(a) for a in net.arch)),
             _row("number of weights", str(net.nw)),
             _row("reference", "synthetic (seeded)"),
             bar, "|                            data information                               |", bar]
    for key in ("train energy data", "test  energy data", "train force  data", "test  force  data",
                "total structures", "max atom number"):
        lines.append(_row(key, str(net.info.get(key, 0))))
    lines.append(_row("standard deviation", "0.000000   eV/atom"))
    lines += [bar, "|                            training details                               |", bar,
              _row("trained on", "synthetic"),
              _row("maise version", version),
              _row("parallelization", "1"),
              _row("cpu and wall times", "0.00  0.00 sec"),
              _row("regularization", "0.0e+00"),
              _row("number of epochs", "0"),
              bar, "|                              performance                                  |", bar]
    for key, unit in (("train energy error", "eV/atom"), ("test  Energy error", "eV/atom"),
                      ("train force  error", "eV/Ang"), ("test  Force  error", "eV/Ang")):
        lines.append(_row(key, f"0.000000  {unit}"))
    lines += [bar, "|                          input component range                            |", bar,
              "        minimum                    maximum                     range"]
    for s in range(S):
        for n in range(D):
            lines.append(f"{net.rmin[s, n]: 21.16f} {net.rmax[s, n]: 25.16f}  {net.drange[s, n]: 25.16f}")
    lines += [bar, "|                        neural network weights                             |", bar]
    for v in net.vector():
        lines.append(f"{v: 21.16f}")
    lines.append(bar)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
        f.write(blocks[0].text())
        for b in blocks[1:]:
            f.write("\n" + b.text())


def read_basis(path: str) -> Basis:
    """Read the basis block (after the ``  B2A`` header) of a model or basis file, unscaled."""
    with open(path) as f:
        lines = f.read().splitlines()
    at = next(i for i, l in enumerate(lines) if l.startswith("  B2A"))
    b2a, scale = (float(v) for v in lines[at + 1].split()[:2])
    rc_row = next(i for i in range(at + 5, len(lines)) if lines[i].startswith(" Rc  "))
    nsym = rc_row - (at + 5) - 1
    funcs = [tuple(int(v) for v in lines[at + 5 + n].split()[1:9]) for n in range(nsym)]
    params = {}
    for k in range(7):
        toks = lines[rc_row + k].split()
        params[toks[0]] = [float(v) for v in toks[2:2 + int(toks[1])]]
    return Basis(funcs=funcs, params=params, b2a=b2a, scale=scale)
