"""Seeded synthetic structures and models for the evaluation workloads.

Workloads (all seeded; sizes as stated):
  * ``cu_fcc``            Cu fcc n x n x n conventional supercell (4 n^3 atoms), Gaussian jitter.
  * ``alloy_fcc``         random substitutional fcc alloy, atoms grouped by species.
  * ``random_triclinic``  random triclinic cell with a minimum pair distance (batched-relaxation inputs).
Model helpers build the synthetic 51-wide (1 species) / species-resolved bases with the reference's
43-function G4 pattern and seeded D-H1-H2-1 tanh networks, written in the ``nn_*`` format.
"""
from __future__ import annotations

import math
import os
from typing import Optional, Sequence, Tuple

import numpy as np

from . import modelfile


def fcc(n: int, a: float) -> Tuple[np.ndarray, np.ndarray]:
    """(lattice[3,3] rows, cartesian positions[4n^3,3]) of an n x n x n cubic fcc supercell."""
    base = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    cells = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 1, 3)
    frac = (cells + base[None]).reshape(-1, 3) / n
    lat = np.eye(3) * (a * n)
    return lat, frac @ lat


def cu_fcc(n: int = 4, a: float = 3.61, sigma: float = 0.05, seed: int = 1):
    lat, x = fcc(n, a)
    rng = np.random.default_rng(seed)
    x = x + rng.normal(0.0, sigma, size=x.shape)
    return np.zeros(len(x), dtype=np.int32), lat, x


def alloy_fcc(n: int, a: float, nspc: int = 2, sigma: float = 0.0, seed: int = 2):
    """Random equal-fraction alloy on an fcc lattice; atoms are reordered so species are grouped."""
    lat, x = fcc(n, a)
    rng = np.random.default_rng(seed)
    N = len(x)
    spc = np.repeat(np.arange(nspc), -(-N // nspc))[:N]
    spc = rng.permutation(spc)
    if sigma > 0:
        x = x + rng.normal(0.0, sigma, size=x.shape)
    order = np.argsort(spc, kind="stable")
    return spc[order].astype(np.int32), lat, x[order]


def random_triclinic(nat: int, nspc: int, rng: np.random.Generator, vol_range=(15.0, 20.0),
                     angle_range=(60.0, 120.0), dmin: float = 2.0, max_tries: int = 2000):
    """Random triclinic cell (volume per atom and angles uniform) with random positions, minimum pair
    distance dmin under periodic boundary conditions, and a composition drawn uniformly on the simplex."""
    while True:
        al, be, ga = np.radians(rng.uniform(*angle_range, size=3))
        ca, cb, cg = math.cos(al), math.cos(be), math.cos(ga)
        g = 1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg
        if g <= 0.05:
            continue
        vol = rng.uniform(*vol_range) * nat
        lens = rng.uniform(0.8, 1.25, size=3)
        scale = (vol / (np.prod(lens) * math.sqrt(g))) ** (1 / 3)
        A, B, Cc = lens * scale
        lat = np.zeros((3, 3))
        lat[0] = [A, 0, 0]
        lat[1] = [B * cg, B * math.sin(ga), 0]
        cx = Cc * cb
        cy = Cc * (ca - cb * cg) / math.sin(ga)
        lat[2] = [cx, cy, math.sqrt(max(Cc * Cc - cx * cx - cy * cy, 1e-12))]
        # reject skewed cells whose plane spacing is below dmin
        rec = np.linalg.inv(lat).T
        if np.min(1.0 / np.linalg.norm(rec, axis=1)) < dmin:
            continue
        # sequential random placement; the periodic distance to every atom placed so far is the minimum over the
        # 27 images around the rounded fractional difference
        shifts = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], dtype=float)
        frac = np.zeros((nat, 3))
        n = 0
        tries = 0
        while n < nat and tries < max_tries * nat:
            tries += 1
            f = rng.random(3)
            if n:
                d = f - frac[:n]
                d -= np.round(d)
                cart = (d[:, None, :] + shifts[None, :, :]) @ lat
                if np.einsum("kiq,kiq->ki", cart, cart).min() < dmin * dmin:
                    continue
            frac[n] = f
            n += 1
        if n < nat:
            continue
        w = rng.dirichlet(np.ones(nspc))
        counts = np.floor(w * nat).astype(int)
        counts[np.argmax(w)] += nat - counts.sum()
        spc = np.repeat(np.arange(nspc), counts).astype(np.int32)
        return spc, lat, frac @ lat


def _cupdag_one(args):
    seq, lo, hi = args
    rng = np.random.default_rng(seq)
    return random_triclinic(int(rng.integers(lo, hi + 1)), 3, rng)


def cupdag_batch(count: int = 4096, seed: int = 3, nat_range=(8, 64), workers: int = 0):
    """The batched-relaxation workload: `count` ternary Cu-Pd-Ag structures with atom counts uniform in
    nat_range, random triclinic cells (15-20 A^3/atom, angles 60-120 deg), random positions with minimum
    periodic pair distance 2.0 A and compositions uniform on the simplex.  Structure k is drawn from its own
    stream (SeedSequence(seed).spawn), so the batch does not depend on how many worker processes build it
    (workers = 0: up to 16 host cores; call before the GPU is initialised in this process)."""
    seqs = np.random.SeedSequence(seed).spawn(count)
    jobs = [(q, nat_range[0], nat_range[1]) for q in seqs]
    if workers == 0:
        workers = min(16, os.cpu_count() or 1)
    if workers <= 1 or count < 64:
        return [_cupdag_one(j) for j in jobs]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_cupdag_one, jobs, chunksize=32)


def synthetic_model(path: str, spcz: Sequence[int], hidden: Tuple[int, int] = (10, 10), seed: int = 7,
                    rc: float = 6.0, range_from: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None,
                    descriptor_max: Optional[np.ndarray] = None, basis=None) -> modelfile.Network:
    """Write a seeded synthetic model: standard basis at cutoff rc (B2A=1, scale=1), D-H1-H2-1 tanh.

    ``descriptor_max`` (shape [S, D]) sets the input ranges to min 0, max = descriptor_max, range = 1/max, so
    the normalised inputs are O(1) and tanh is not saturated.  Without it the ranges are (0, 1, 1).
    ``basis``: a Basis or a list of one per species in place of the standard one.
    """
    if basis is None:
        basis = modelfile.standard_basis(rc=rc)
    S = len(spcz)
    D = (basis[0] if isinstance(basis, (list, tuple)) else basis).width(S)
    net = modelfile.random_network(spcz, [D, hidden[0], hidden[1], 1], seed)
    if descriptor_max is not None:
        mx = np.maximum(np.asarray(descriptor_max, dtype=np.float64).reshape(S, D), 1e-6)
        net.rmax = mx
        net.drange = 1.0 / mx
    modelfile.write_model(path, net, basis)
    return net


def calibrated_model(path: str, spcz, hidden, seed, sample_structures, descriptor_fn, rc: float = 6.0, basis=None):
    """Two-step synthetic model: write with unit ranges, compute descriptors of the sample structures with
    ``descriptor_fn(model_path, spc, lat, x) -> G[N, D]``, then rewrite with ranges 1/max(G) per species."""
    synthetic_model(path, spcz, hidden, seed, rc, basis=basis)
    S = len(spcz)
    b0 = modelfile.standard_basis(rc=rc) if basis is None else (basis[0] if isinstance(basis, (list, tuple)) else basis)
    first = b0.width(S)
    gmax = np.zeros((S, first))
    for spc, lat, x in sample_structures:
        G = descriptor_fn(path, spc, lat, x)
        for s in range(S):
            sel = spc == s
            if sel.any():
                gmax[s] = np.maximum(gmax[s], G[sel].max(axis=0))
    gmax[gmax <= 0] = 1.0
    synthetic_model(path, spcz, hidden, seed, rc, descriptor_max=gmax, basis=basis)
    return path
