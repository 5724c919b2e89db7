"""Evolutionary structure search with the learned screening surrogate.

The generation loop of the reference's EVOS caller (src/emod.c:484-757) with its bulk operators:
  crossover  cut-and-splice after BULK_MATE (src/efnc.c:2021-2310): each parent is shifted by a random
             fractional vector, the first parent is cut by a random plane through a point near the cell centre,
             the second parent's complementary slab is taken with the plane offset stepped until the atom count
             and composition come out right; the child keeps the lattice of one parent (MATE_LT,
             src/eutl.c:139-172: the second parent's lattice, cyclically permuted to best match the first, with
             probability equal to the first parent's share of atoms) and is rejected when its volume differs by
             10% or more from the parents' mean;
  mutation   Gaussian displacement of every atom (sigma in A) of each crossover child, then the minimum-distance
             check; and, for a `mute_frac` share of the offspring, BULK_MUTE's operator (src/efnc.c:2316-2386):
             a parent drawn by weight, distorted by SHKE_CL (src/eutl.c:176-218: random symmetric lattice strain,
             atomic displacements), reduced (SHRT_LV), with species swapped at rate `mute_cs`;
  selection  the population is ranked by relaxed energy per atom (RANK_TR, src/emod.c:304-444); a structure whose
             COMP_CL similarity (src/eutl.c:394-413: 0.8 CxC of the RDF fingerprints of src/cell.c:369-427 + 0.1 of
             an energy and 0.1 of a volume closeness) to a lower-energy member exceeds `dup_cut` is a duplicate and
             dropped, and parents are drawn with weights that fall off with rank.
All structures are relaxed on the GPU in one batch (mg_relax, positions only, cell fixed).

Surrogate screening: each generation's offspring are evaluated once unrelaxed (one batched descriptor pass),
the surrogate predicts their relaxed energy per atom from the per-species means of their descriptor rows, and
only the predicted best `keep_top` fraction plus a random `keep_random` fraction of the rest are relaxed.  The
surrogate is retrained after every generation on all (unrelaxed features, relaxed energy) pairs relaxed so far.
With `screen=False` every offspring is relaxed and no surrogate is trained (the method without it).  With
`audit=True` every offspring is also relaxed (separately timed, outside the search clock) so that top-k recall
and the surrogate's energy error are measured against full relaxation.

Relaxation tail: the reference gives each relaxation job a wall-time limit and drops the ones that exceed it
(RELX_TR, src/emod.c:608-660).  Here a generation's batch runs at most `relax_cap` FIRE steps; a structure not
converged by then is carried, at its current positions, into the next generation's batch (FIRE restarts from
rest) and joins the population when it converges; one still unconverged after `max_steps` steps in total is
dropped, as the reference drops a job over its limit.  So the slowest few structures no longer set every
generation's wall time.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import Context, Model, _check, _dp, _fp, _ip, lib
from .surrogate import Surrogate

Structure = Tuple[np.ndarray, np.ndarray, np.ndarray]  # (species[N], lattice rows[3,3], Cartesian x[N,3])


@dataclass
class GAConfig:
    pop: int = 512
    generations: int = 30
    counts: Tuple[int, ...] = (16, 8)          # atoms per species (Mg16Ca8: Mg:Ca = 2:1)
    vol_range: Tuple[float, float] = (24.0, 32.0)  # A^3 per atom of the initial random cells
    sigma: float = 0.3                         # Gaussian mutation (A)
    dmin: float = 2.0                          # minimum interatomic distance of a new structure (A)
    keep_top: float = 0.25
    keep_random: float = 0.10
    fmax: float = 0.01
    max_steps: int = 1500
    urep: float = 100.0
    seed: int = 5
    train_steps: int = 300                     # Adam steps per generation (warm start from the last generation)
    batch: int = 64
    lr: float = 1e-3
    audit: bool = True                         # also relax every offspring to measure recall and error
    mute_frac: float = 0.0                     # share of offspring made by BULK_MUTE-style mutation (rest: crossover)
    mute_cl: float = 0.15                      # lattice distortion of the mutation (LDST)
    mute_ca: float = 0.15                      # atomic displacement of the mutation, A (ADST)
    mute_cs: float = 0.10                      # species swap rate of the mutation (SDST)
    dup_cut: float = 0.95                      # COMP_CL above this marks a duplicate (SCUT); >= 1 disables the test
    fp_nrdf: int = 3000                        # fingerprint: RDF bins (NRDF), up to fp_rmax (RMAX) with a cosine
    fp_rmax: float = 6.0                       # taper from fp_rmin (RMIN), Gaussian spread fp_sigma (DELR), A
    fp_rmin: float = 5.8
    fp_sigma: float = 0.008
    precision: int = 0                         # MLP precision of the relaxations and evaluations (MG_PREC_*)
    screen: bool = True                        # screen offspring with the surrogate (False: relax every offspring)
    offspring: int = 0                         # offspring per generation (0: pop)
    relax_cap: int = 0                         # FIRE steps per generation before a structure is carried over (0: none)
    optimizer: str = "fire"                    # relaxation method: "fire" (mg_relax) or "lbfgs" (mg_relax_lbfgs)
    prerelax: int = 0                          # screening: relaxation steps of every offspring before it is ranked


@dataclass
class GenerationReport:
    gen: int
    offspring: int
    relaxed: int
    seconds_screened: float                    # the whole generation: offspring, screening, relaxation, training, update
    seconds_full: Optional[float] = None       # relaxation of every offspring (audit)
    recall_top10: Optional[float] = None
    recall_top25: Optional[float] = None
    mae_eV_atom: Optional[float] = None
    best_eV_atom: float = math.nan
    train_loss: Optional[float] = None
    seconds_offspring: Optional[float] = None  # crossover and mutation (host)
    seconds_prerelax: Optional[float] = None   # the first prerelax steps of every offspring (cfg.prerelax)
    seconds_predict: Optional[float] = None    # unrelaxed evaluation + features + prediction + selection
    seconds_relax: Optional[float] = None      # relaxation of the kept offspring
    seconds_train: Optional[float] = None      # retraining: host time to start it plus any wait for it after the update
    seconds_select: Optional[float] = None     # fingerprints and the population update
    elapsed: float = 0.0                       # search clock at the end of the generation (audit excluded)
    carried: int = 0                           # unconverged structures carried into the next generation
    dropped: int = 0                           # structures dropped after max_steps FIRE steps in total


@dataclass
class GAResult:
    reports: List[GenerationReport] = field(default_factory=list)
    population_E: Optional[np.ndarray] = None
    best: Optional[Structure] = None


# ------------------------------------------------------------------ geometry helpers
def lattice_params(lat: np.ndarray):
    a, b, c = (np.linalg.norm(v) for v in lat)
    al = math.acos(np.dot(lat[1], lat[2]) / (b * c))
    be = math.acos(np.dot(lat[0], lat[2]) / (a * c))
    ga = math.acos(np.dot(lat[0], lat[1]) / (a * b))
    return np.array([a, b, c]), np.array([al, be, ga])


def lattice_from_params(L, A) -> np.ndarray:
    a, b, c = L
    al, be, ga = A
    ca, cb, cg, sg = math.cos(al), math.cos(be), math.cos(ga), math.sin(ga)
    cx = c * cb
    cy = c * (ca - cb * cg) / sg
    cz2 = c * c - cx * cx - cy * cy
    if cz2 <= 1e-6:
        return None
    return np.array([[a, 0.0, 0.0], [b * cg, b * sg, 0.0], [cx, cy, math.sqrt(cz2)]])


def min_distance(lat: np.ndarray, x: np.ndarray) -> float:
    """Smallest periodic interatomic distance (native, mg_evo_min_distance)."""
    lat = np.ascontiguousarray(lat, dtype=np.float64).reshape(9)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    return float(lib().mg_evo_min_distance(_dp(lat), _dp(x), len(x) // 3))


def random_structure(cfg: GAConfig, rng: np.random.Generator) -> Structure:
    from .workloads import random_triclinic
    n = sum(cfg.counts)
    _, lat, x = random_triclinic(n, 1, rng, vol_range=cfg.vol_range, dmin=cfg.dmin)
    spc = np.repeat(np.arange(len(cfg.counts)), cfg.counts).astype(np.int32)
    return spc, lat, x


# ------------------------------------------------------------------ operators (native: mg_evo_offspring)
def make_offspring(parents: List[Structure], weights: np.ndarray, cfg: GAConfig, rng: np.random.Generator,
                   count: int, max_tries: int = 500) -> List[Structure]:
    """count children by crossover + Gaussian mutation + minimum-distance test; a slot whose operators keep
    failing gets a fresh random structure."""
    N = len(parents[0][0])
    spc = np.ascontiguousarray(np.concatenate([p[0] for p in parents]), dtype=np.int32)
    lat = np.ascontiguousarray(np.concatenate([np.asarray(p[1]).reshape(9) for p in parents]), dtype=np.float64)
    x = np.ascontiguousarray(np.concatenate([np.asarray(p[2]).reshape(-1) for p in parents]), dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    os_ = np.zeros(count * N, dtype=np.int32)
    ol = np.zeros(count * 9)
    ox = np.zeros(count * 3 * N)
    failed = np.zeros(count, dtype=np.int32)
    nspc = len(cfg.counts)
    rc = lib().mg_evo_offspring(len(parents), N, nspc, _ip(spc), _dp(lat), _dp(x), _dp(w), count, float(cfg.sigma),
                                float(cfg.dmin), int(rng.integers(1 << 62)), int(max_tries), _ip(os_), _dp(ol), _dp(ox),
                                _ip(failed))
    if rc < 0:
        _check(rc, "mg_evo_offspring")
    out: List[Structure] = []
    for k in range(count):
        if failed[k]:
            out.append(random_structure(cfg, rng))
        else:
            out.append((os_[k * N:(k + 1) * N].copy(), ol[9 * k:9 * k + 9].reshape(3, 3).copy(),
                        ox[3 * N * k:3 * N * (k + 1)].reshape(N, 3).copy()))
    return out


def mutate(parents: List[Structure], weights: np.ndarray, cfg: GAConfig, rng: np.random.Generator, count: int,
           max_tries: int = 500) -> List[Structure]:
    """count mutants (native mg_evo_mutate, after BULK_MUTE + SHKE_CL); a slot whose tries all fail gets a fresh
    random structure."""
    N = len(parents[0][0])
    spc = np.ascontiguousarray(np.concatenate([p[0] for p in parents]), dtype=np.int32)
    lat = np.ascontiguousarray(np.concatenate([np.asarray(p[1]).reshape(9) for p in parents]), dtype=np.float64)
    x = np.ascontiguousarray(np.concatenate([np.asarray(p[2]).reshape(-1) for p in parents]), dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    os_ = np.zeros(count * N, dtype=np.int32)
    ol = np.zeros(count * 9)
    ox = np.zeros(count * 3 * N)
    failed = np.zeros(count, dtype=np.int32)
    rc = lib().mg_evo_mutate(len(parents), N, len(cfg.counts), _ip(spc), _dp(lat), _dp(x), _dp(w), count,
                             float(cfg.mute_cl), float(cfg.mute_ca), float(cfg.mute_cs), float(cfg.dmin),
                             int(rng.integers(1 << 62)), int(max_tries), _ip(os_), _dp(ol), _dp(ox), _ip(failed))
    if rc < 0:
        _check(rc, "mg_evo_mutate")
    out: List[Structure] = []
    for k in range(count):
        if failed[k]:
            out.append(random_structure(cfg, rng))
        else:
            out.append((os_[k * N:(k + 1) * N].copy(), ol[9 * k:9 * k + 9].reshape(3, 3).copy(),
                        ox[3 * N * k:3 * N * (k + 1)].reshape(N, 3).copy()))
    return out


def fingerprints(structs: List[Structure], nspc: int, cfg: Optional[GAConfig] = None) -> np.ndarray:
    """Unit-length RDF fingerprints [S, nspc(nspc+1)/2 * nrdf] (native mg_evo_fingerprint, after RDF(C, 1))."""
    cfg = cfg or GAConfig()
    S = len(structs)
    npair = nspc * (nspc + 1) // 2
    off = np.zeros(S + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(s[0]) for s in structs])
    spc = np.ascontiguousarray(np.concatenate([s[0] for s in structs]), dtype=np.int32)
    lat = np.ascontiguousarray(np.concatenate([np.asarray(s[1]).reshape(9) for s in structs]), dtype=np.float64)
    x = np.ascontiguousarray(np.concatenate([np.asarray(s[2]).reshape(-1) for s in structs]), dtype=np.float64)
    out = np.zeros((S, npair * cfg.fp_nrdf), dtype=np.float32)
    _check(lib().mg_evo_fingerprint(S, _ip(off), nspc, _ip(spc), _dp(lat), _dp(x), int(cfg.fp_nrdf),
                                    float(cfg.fp_rmax), float(cfg.fp_rmin), float(cfg.fp_sigma), _fp(out)),
           "mg_evo_fingerprint")
    return out


def cell_volumes(structs: List[Structure]) -> np.ndarray:
    return np.array([abs(np.linalg.det(np.asarray(s[1]).reshape(3, 3))) for s in structs])


def comp_cl(fa: np.ndarray, fb: np.ndarray, ea, eb, va, vb) -> np.ndarray:
    """COMP_CL (src/eutl.c:394-413) of structure a against the structures b: 0.8 CxC (fingerprint dot product)
    + 0.1 exp(-dE^2 / (2 0.05^2)) + 0.1 exp(-dV^2 / (2 0.01^2)), dV = (V_a - V_b) / (2 (V_a + V_b)); at most 1."""
    fb = np.atleast_2d(fb)
    cxc = fb @ fa
    de = np.asarray(eb, dtype=np.float64) - float(ea)
    dv = 0.5 * (float(va) - np.asarray(vb, dtype=np.float64)) / (float(va) + np.asarray(vb, dtype=np.float64))
    return 0.8 * cxc + 0.1 * np.exp(-0.5 * de * de / 0.05 ** 2) + 0.1 * np.exp(-0.5 * dv * dv / 0.01 ** 2)


# COMP_CL = 0.8 CxC + 0.1 e_E + 0.1 e_V with CxC, e_E, e_V <= 1 exceeds 0.95 only if e_E > 0.05 and e_V > 0.05, i.e.
# |dE| < 0.05 sqrt(2 ln 20) eV/atom and |dV| < 0.01 sqrt(2 ln 20): the only pairs whose fingerprints need comparing
_DUP_DE = 0.05 * math.sqrt(2.0 * math.log(20.0))
_DUP_DV = 0.01 * math.sqrt(2.0 * math.log(20.0))


def unique_by_fingerprint(E: np.ndarray, fp: np.ndarray, V: np.ndarray, cut: float, limit: int,
                          n_unique: int = 0) -> List[int]:
    """Indices of up to `limit` structures in increasing energy, skipping every structure whose COMP_CL with one
    already kept exceeds `cut` (RANK_TR's duplicate test, src/emod.c:351-366, which divides such a structure's
    fitness by ~1000; here it leaves the population).  Infinite energies are dropped.  The first `n_unique`
    structures are known to be no duplicates of each other (the previous population), so only pairs with at least one
    later structure are compared: their fingerprint products come from one matrix product, and the walk itself
    (native mg_evo_unique) scores only pairs inside the energy and volume windows where COMP_CL can exceed 0.95
    (exact for cut >= 0.95)."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    n = len(E)
    test = cut < 1.0 and n > n_unique
    G = np.ascontiguousarray(fp[n_unique:] @ fp.T, dtype=np.float32) if test else np.zeros(1, np.float32)
    de_win = _DUP_DE if cut >= 0.95 else np.inf
    dv_win = _DUP_DV if cut >= 0.95 else np.inf
    keep = np.zeros(max(n, 1), dtype=np.int32)
    nk = lib().mg_evo_unique(n, _dp(E), _dp(V), _fp(G), int(n_unique), float(cut), float(de_win), float(dv_win),
                             int(limit), _ip(keep))
    if nk < 0:
        _check(nk, "mg_evo_unique")
    return [int(k) for k in keep[:nk]]


# ------------------------------------------------------------------ the search
class SurrogateGA:
    def __init__(self, model_path: str, cfg: GAConfig, device: int = 0):
        self.cfg = cfg
        self.model = Model(model_path)
        self.device = device
        self.rng = np.random.default_rng(cfg.seed)
        self.nin = self.model.nspc * self.model.D
        self.sur = Surrogate(self.nin, device=device, seed=cfg.seed)
        self.trained = False
        self.X: List[np.ndarray] = []
        self.y: List[np.ndarray] = []
        self._ctx_eval = Context(self.model, device=device, urep=cfg.urep, precision=cfg.precision)
        self._ctx_relax = Context(self.model, device=device, urep=cfg.urep, precision=cfg.precision)

    def close(self):
        self.sur.close()
        self._ctx_eval.close()
        self._ctx_relax.close()

    def features(self, structs: List[Structure]) -> np.ndarray:
        c = self._ctx_eval
        c.set_batch(structs)
        c.eval(forces=False)
        return c.features()

    def relax(self, structs: List[Structure], max_steps: Optional[int] = None):
        """Relax in one batch with cfg.optimizer; returns (structures at their final positions, energy per atom,
        converged flags, steps taken)."""
        c = self._ctx_relax
        c.set_batch(structs)
        ms = self.cfg.max_steps if max_steps is None else max_steps
        if self.cfg.optimizer == "lbfgs":
            steps, done = c.relax_lbfgs(max_steps=ms, fmax=self.cfg.fmax)
        elif self.cfg.optimizer == "fire":
            steps, done = c.relax(max_steps=ms, fmax=self.cfg.fmax)
        else:
            raise ValueError(f"unknown optimizer {self.cfg.optimizer!r}")
        E = c.batch_result(forces=False)["E"]
        Xall = c.batch_positions()
        out, epa = [], np.zeros(len(structs))
        a = 0
        for k, (spc, lat, _) in enumerate(structs):
            n = len(spc)
            out.append((spc, lat, Xall[a:a + n]))
            epa[k] = E[k] / n
            a += n
        return out, epa, done, steps

    def train(self) -> float:
        self.train_begin()
        return self.train_end()

    def train_end(self) -> float:
        """Wait for the training run started by train_begin; returns its final loss."""
        return self.sur.train_end()

    def train_begin(self) -> None:
        """A fixed budget of train_steps Adam steps, warm-started from the last generation's network, over the
        newest generation's pairs plus a uniform sample of the earlier ones (all of them while they fit in the
        budget's train_steps * batch samples, so that the cost per generation stays flat as the data grows); the
        normalisation is fitted once, on the first (random, fully relaxed) generation.  The run is left in flight
        on the device (the host goes on with the population update meanwhile)."""
        cap = self.cfg.train_steps * self.cfg.batch
        X_new, y_new = self.X[-1], self.y[-1]
        if len(self.X) > 1:
            X_old = np.concatenate(self.X[:-1])
            y_old = np.concatenate(self.y[:-1])
            room = cap - len(y_new)
            if len(y_old) > room:
                pick = self.rng.choice(len(y_old), size=max(room, 0), replace=False)
                X_old, y_old = X_old[pick], y_old[pick]
            X = np.concatenate([X_old, X_new])
            y = np.concatenate([y_old, y_new])
        else:
            X, y = X_new, y_new
        if not self.trained:
            self.sur.fit_norm(X, y)
        epochs = max(1, int(round(cap / len(y))))
        self.sur.train_begin(X, y, epochs=epochs, batch=self.cfg.batch, lr=self.cfg.lr,
                             seed=int(self.rng.integers(1 << 62)))
        self.trained = True

    def select(self, pred: np.ndarray) -> np.ndarray:
        n = len(pred)
        ntop = int(round(self.cfg.keep_top * n))
        nrand = int(round(self.cfg.keep_random * n))
        order = np.argsort(pred, kind="stable")
        top = order[:ntop]
        rest = order[ntop:]
        extra = self.rng.choice(rest, size=min(nrand, len(rest)), replace=False) if nrand else np.array([], int)
        return np.sort(np.concatenate([top, extra]).astype(int))

    def update_population(self, pop: List[Structure], popE: np.ndarray, popFP: Optional[np.ndarray],
                          new: List[Structure], newE: np.ndarray):
        """Keep the cfg.pop lowest-energy structures of pop + new that are not duplicates (COMP_CL > dup_cut) of a
        lower-energy one; returns (structures, energies, fingerprints)."""
        allS = pop + new
        allE = np.concatenate([popE, newE])
        newFP = fingerprints(new, len(self.cfg.counts), self.cfg) if new else None
        parts = [f for f in (popFP, newFP) if f is not None and len(f)]
        allFP = np.concatenate(parts) if parts else np.zeros((0, 1), np.float32)
        keep = unique_by_fingerprint(allE, allFP, cell_volumes(allS), self.cfg.dup_cut, self.cfg.pop,
                                     n_unique=len(pop))
        return [allS[k] for k in keep], allE[keep], allFP[keep]

    def _relax_with_carry(self, batch, feats, prior_steps, capped=True):
        """Relax `batch` for at most relax_cap steps (max_steps without a cap, or with capped=False).  Returns the
        converged structures with their energies and features, and the carry list of (structure, features, steps
        so far) for the unconverged ones still under max_steps, plus the number dropped."""
        cfg = self.cfg
        cap = cfg.relax_cap if (cfg.relax_cap > 0 and capped) else cfg.max_steps
        rel, epa, done, steps = self.relax(batch, max_steps=cap)
        fin, finE, finF, carry, dropped = [], [], [], [], 0
        for k in range(len(batch)):
            total = prior_steps[k] + int(steps[k])
            if done[k]:
                fin.append(rel[k])
                finE.append(epa[k])
                finF.append(feats[k])
            elif total < cfg.max_steps and cfg.relax_cap > 0:
                carry.append((rel[k], feats[k], total))
            else:
                dropped += 1
        return fin, np.array(finE), finF, carry, dropped

    def run(self, log=None, time_budget: Optional[float] = None) -> GAResult:
        """Run up to cfg.generations generations, or until the search clock passes time_budget seconds."""
        cfg = self.cfg
        res = GAResult()
        screen = cfg.screen
        n_off = cfg.offspring if cfg.offspring > 0 else cfg.pop
        clock = 0.0
        # generation 0: random structures, all relaxed without the cap (they are the first parents), first
        # training set
        t0 = time.perf_counter()
        init = [random_structure(cfg, self.rng) for _ in range(cfg.pop)]
        feat = self.features(init) if screen else [None] * len(init)
        fin, finE, finF, carry, dropped = self._relax_with_carry(init, feat, [0] * len(init), capped=False)
        loss = None
        if screen and fin:
            self.X.append(np.stack(finF))
            self.y.append(finE.astype(np.float32))
            loss = self.train()
        pop, popE, popFP = self.update_population([], np.zeros(0), None, fin, finE)
        clock += time.perf_counter() - t0
        res.reports.append(GenerationReport(0, len(init), len(init), clock, best_eV_atom=float(popE[0]),
                                            train_loss=loss, elapsed=clock, carried=len(carry), dropped=dropped))
        if log:
            log(res.reports[-1])
        for g in range(1, cfg.generations + 1):
            if time_budget is not None and clock >= time_budget:
                break
            t0 = time.perf_counter()
            ranks = np.arange(len(pop))
            weights = np.exp(-ranks / max(1.0, 0.25 * len(pop)))
            nm = int(round(cfg.mute_frac * n_off))
            kids = make_offspring(pop, weights, cfg, self.rng, n_off - nm)
            if nm:
                kids += mutate(pop, weights, cfg, self.rng, nm)
            t_off = time.perf_counter()
            pred = None
            ranked, pre_steps = kids, [0] * len(kids)
            if screen and cfg.prerelax > 0:
                # every offspring takes a few relaxation steps first, in one batch: the surrogate then ranks
                # structures that have begun to settle (and is trained on such structures' features), and the kept
                # ones go on from there
                ranked, _, _, st_pre = self.relax(kids, max_steps=cfg.prerelax)
                pre_steps = [int(v) for v in st_pre]
            t_pre = time.perf_counter()
            if screen:
                feat = self.features(ranked)
                pred = self.sur.predict(feat)
                sel = self.select(pred)
            else:
                feat = [None] * len(kids)
                sel = np.arange(len(kids))
            t1 = time.perf_counter()
            batch = [ranked[k] for k in sel] + [c[0] for c in carry]
            bfeat = [feat[k] for k in sel] + [c[1] for c in carry]
            prior = [pre_steps[k] for k in sel] + [c[2] for c in carry]
            fin, finE, finF, carry, dropped = self._relax_with_carry(batch, bfeat, prior)
            t2 = time.perf_counter()
            loss = None
            training = False
            if screen and fin:
                self.X.append(np.stack(finF))
                self.y.append(finE.astype(np.float32))
                self.train_begin()             # runs on the device during the population update
                training = True
            t3 = time.perf_counter()
            pop, popE, popFP = self.update_population(pop, popE, popFP, fin, finE)
            t4 = time.perf_counter()
            if training:
                loss = self.train_end()
            t5 = time.perf_counter()
            clock += t5 - t0
            rep = GenerationReport(g, len(kids), len(sel), t5 - t0, train_loss=loss, seconds_predict=t1 - t0,
                                   seconds_relax=t2 - t1, seconds_train=(t3 - t2) + (t5 - t4), elapsed=clock,
                                   carried=len(carry), dropped=dropped)
            rep.seconds_select = t4 - t3
            rep.seconds_offspring = t_off - t0
            rep.seconds_prerelax = t_pre - t_off
            rep.seconds_predict = t1 - t_pre
            if screen and cfg.audit:
                t1 = time.perf_counter()
                _, epa_all, done_all, _ = self.relax(kids)
                rep.seconds_full = time.perf_counter() - t1
                order = np.argsort(epa_all, kind="stable")
                sset = set(sel.tolist())
                for frac_k, name in ((0.10, "recall_top10"), (0.25, "recall_top25")):
                    k = max(1, int(round(frac_k * len(kids))))
                    setattr(rep, name, len(sset.intersection(order[:k].tolist())) / k)
                rep.mae_eV_atom = float(np.mean(np.abs(pred.astype(np.float64) - epa_all)))
            rep.best_eV_atom = float(popE[0])
            res.reports.append(rep)
            if log:
                log(rep)
        res.population_E = popE
        res.best = pop[0]
        return res


def compare_screening(model_path: str, cfg: GAConfig, seconds: float, seeds, variants=None, log=None) -> dict:
    """The same seeded search with and without surrogate screening, each run for `seconds` of search clock.

    variants maps a name to GAConfig overrides; the default compares the method without the surrogate
    ("unscreened": every offspring relaxed), the screened search as configured ("screened") and the screened
    search drawing as many offspring as make its relaxed share equal the unscreened batch ("screened_wide").
    For each variant and seed the result holds the best energy per atom against the search clock; per seed the
    target is the lowest energy any variant reached (within 1e-4 eV/atom), and `seconds_to_target` the clock at
    which a variant first reached it (None if it did not)."""
    import dataclasses
    share = cfg.keep_top + cfg.keep_random
    if variants is None:
        variants = {"unscreened": {"screen": False},
                    "screened": {"screen": True},
                    "screened_wide": {"screen": True, "offspring": int(round(cfg.pop / share))}}
    runs = {name: [] for name in variants}
    for seed in seeds:
        for name, over in variants.items():
            c = dataclasses.replace(cfg, seed=int(seed), audit=False, generations=10 ** 6, **over)
            ga = SurrogateGA(model_path, c)
            try:
                res = ga.run(time_budget=seconds)
            finally:
                ga.close()
            traj = [(r.elapsed, r.best_eV_atom) for r in res.reports]
            runs[name].append({"seed": int(seed), "generations": len(res.reports) - 1,
                               "relaxed_per_generation": res.reports[-1].relaxed if len(res.reports) > 1 else None,
                               "best_eV_atom": float(res.population_E[0]), "trajectory": traj})
            if log:
                log(name, seed, runs[name][-1])
    for k, seed in enumerate(seeds):
        target = min(runs[name][k]["best_eV_atom"] for name in variants)
        for name in variants:
            hit = [t for t, e in runs[name][k]["trajectory"] if e <= target + 1e-4]
            runs[name][k]["target_eV_atom"] = target
            runs[name][k]["seconds_to_target"] = hit[0] if hit else None
    summary = {}
    for name in variants:
        r = runs[name]
        tt = [x["seconds_to_target"] for x in r]
        summary[name] = {"best_eV_atom_mean": float(np.mean([x["best_eV_atom"] for x in r])),
                         "reached_target": int(sum(t is not None for t in tt)),
                         "seconds_to_target_median": (float(np.median([t if t is not None else np.inf for t in tt]))
                                                      if tt else None),
                         "generations_mean": float(np.mean([x["generations"] for x in r]))}
    return {"seconds": seconds, "seeds": [int(s) for s in seeds], "summary": summary, "runs": runs}
