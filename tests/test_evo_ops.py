"""Host operators of the evolutionary search (no GPU): the BULK_MUTE-style mutation (src/efnc.c:2316-2386 with SHKE_CL,
src/eutl.c:176-218), the RDF structure fingerprint (RDF(C, 1), src/cell.c:369-415) with CxC and COMP_CL
(src/cell.c:417-427, src/eutl.c:394-413), and the duplicate filter of the population update (RANK_TR,
src/emod.c:351-366).  Checked by invariants and by the fingerprint's definition evaluated term by term on a tiny
cell (the native code uses a recurrence across the bins)."""
import math

import numpy as np
import pytest

from maise_amd import evolve as ev
from maise_amd import workloads as wl


def _parents(n=4, seed=3):
    cfg = ev.GAConfig(counts=(6, 3))
    rng = np.random.default_rng(seed)
    return cfg, [ev.random_structure(cfg, rng) for _ in range(n)]


def test_mutation_keeps_composition_grouping_and_distances():
    cfg, par = _parents()
    kids = ev.mutate(par, np.ones(len(par)), cfg, np.random.default_rng(1), 40)
    assert len(kids) == 40
    for spc, lat, x in kids:
        np.testing.assert_array_equal(spc, par[0][0])           # species grouped, counts kept
        assert ev.min_distance(lat, x) >= cfg.dmin
        V = abs(np.linalg.det(lat))
        Vp = [abs(np.linalg.det(p[1])) for p in par]
        assert min(Vp) * 0.5 <= V <= max(Vp) * 1.5               # SHKE_CL's volume window
        # reduced rows (SHRT_LV): no row projects onto another by more than half its squared length
        for a in range(3):
            for b in range(3):
                if a != b:
                    assert abs(lat[a] @ lat[b]) <= 0.5 * (lat[b] @ lat[b]) * (1 + 1e-9)


def test_mutation_is_seeded():
    cfg, par = _parents()
    a = ev.mutate(par, np.ones(len(par)), cfg, np.random.default_rng(9), 8)
    b = ev.mutate(par, np.ones(len(par)), cfg, np.random.default_rng(9), 8)
    for (s1, l1, x1), (s2, l2, x2) in zip(a, b):
        np.testing.assert_array_equal(l1, l2)
        np.testing.assert_array_equal(x1, x2)


def test_mutation_without_distortion_is_the_parent_and_swaps_permute_positions():
    cfg, par = _parents(1)
    # a reduced parent lattice, no strain, no displacement, no swap: the child is the parent
    spc, lat, x = par[0]
    cfg0 = ev.GAConfig(counts=(6, 3), mute_cl=0.0, mute_ca=0.0, mute_cs=0.0, dmin=0.5)
    k = ev.mutate(par, np.ones(1), cfg0, np.random.default_rng(2), 1)[0]
    np.testing.assert_allclose(k[1], lat, atol=1e-12)
    np.testing.assert_allclose(k[2], x, atol=1e-10)
    # swaps only: the same multiset of positions, and at least one atom of each species has moved
    cfg1 = ev.GAConfig(counts=(6, 3), mute_cl=0.0, mute_ca=0.0, mute_cs=1.0, dmin=0.5)
    k = ev.mutate(par, np.ones(1), cfg1, np.random.default_rng(3), 1)[0]
    key = lambda a: a[np.lexsort(np.round(a, 8).T)]  # noqa: E731
    np.testing.assert_allclose(key(k[2]), key(x), atol=1e-10)
    assert not np.allclose(k[2], x)


def _rdf_by_definition(spc, lat, x, nspc, nrdf, rmax, rmin, sigma):
    """RDF(C, 1) evaluated bin by bin with an explicit exp per term (tiny cells only)."""
    dr = rmax / nrdf
    Do = int(3 * sigma / dr)
    pidx = {}
    p = 0
    for m in range(nspc):
        for n in range(m, nspc):
            pidx[(m, n)] = p
            p += 1
    acc = np.zeros((p, nrdf))
    rcut = rmax + 3 * sigma + dr
    K = 3
    for i in range(len(spc)):
        for j in range(len(spc)):
            if spc[j] < spc[i]:
                continue
            for a in range(-K, K + 1):
                for b in range(-K, K + 1):
                    for c in range(-K, K + 1):
                        if i == j and a == b == c == 0:
                            continue
                        d = x[j] - x[i] + a * lat[0] + b * lat[1] + c * lat[2]
                        r = math.sqrt(d @ d)
                        if r >= rcut:
                            continue
                        o = int(r / dr)
                        for k in range(o - Do, o + Do + 1):
                            if 0 <= k < nrdf:
                                xk = k * dr
                                f = 1.0 if xk < rmin else math.cos(0.5 * math.pi * (xk - rmin) / (rmax - rmin))
                                acc[pidx[(spc[i], spc[j])], k] += f * math.exp(-(r - xk) ** 2 * 0.5 / sigma ** 2)
    for m in range(nspc):
        acc[pidx[(m, m)]] *= 0.5
    return (acc / np.sqrt((acc ** 2).sum())).ravel()


def test_fingerprint_matches_its_definition_on_a_tiny_cell():
    rng = np.random.default_rng(4)
    lat = np.array([[3.1, 0.2, 0.0], [0.3, 3.3, 0.1], [0.0, 0.4, 3.2]])
    x = rng.random((3, 3)) @ lat
    spc = np.array([0, 0, 1], np.int32)
    cfg = ev.GAConfig(counts=(2, 1), fp_nrdf=300)
    fp = ev.fingerprints([(spc, lat, x)], 2, cfg)[0]
    ref = _rdf_by_definition(spc, lat, x, 2, 300, cfg.fp_rmax, cfg.fp_rmin, cfg.fp_sigma * 10)
    # the definition with a 10x wider spread, to check the recurrence across many bins as well
    cfgw = ev.GAConfig(counts=(2, 1), fp_nrdf=300, fp_sigma=cfg.fp_sigma * 10)
    fpw = ev.fingerprints([(spc, lat, x)], 2, cfgw)[0]
    np.testing.assert_allclose(fpw, ref, atol=1e-6)
    ref0 = _rdf_by_definition(spc, lat, x, 2, 300, cfg.fp_rmax, cfg.fp_rmin, cfg.fp_sigma)
    np.testing.assert_allclose(fp, ref0, atol=1e-6)


def test_fingerprint_invariances_and_similarity():
    cfg = ev.GAConfig(counts=(16, 8))
    rng = np.random.default_rng(6)
    s = ev.random_structure(cfg, rng)
    spc, lat, x = s
    fp = ev.fingerprints([s], 2, cfg)[0]
    assert abs(np.linalg.norm(fp) - 1.0) < 1e-5
    # translation, permutation within a species, wrapping by a lattice vector: the same fingerprint
    perm = np.concatenate([rng.permutation(16), 16 + rng.permutation(8)])
    x2 = x[perm] + np.array([0.7, -1.3, 2.1])
    x2[3] += lat[1]
    fp2 = ev.fingerprints([(spc, lat, x2)], 2, cfg)[0]
    assert fp @ fp2 > 1 - 1e-5
    # a 2x1x1 supercell: every distance distribution doubles, the unit-length fingerprint is the same
    spc_s = np.concatenate([spc[:16], spc[:16], spc[16:], spc[16:]])
    x_s = np.concatenate([x[:16], x[:16] + lat[0], x[16:], x[16:] + lat[0]])
    lat_s = lat.copy()
    lat_s[0] *= 2
    fp_s = ev.fingerprints([(spc_s, lat_s, x_s)], 2, cfg)[0]
    assert fp @ fp_s > 1 - 1e-4
    # another random structure is not a duplicate
    o = ev.random_structure(cfg, rng)
    fpo = ev.fingerprints([o], 2, cfg)[0]
    assert fp @ fpo < 0.9
    # COMP_CL is at most 1, and exactly 1 for identical structure, energy and volume
    V = abs(np.linalg.det(lat))
    assert ev.comp_cl(fp, fp[None], -1.5, [-1.5], V, [V])[0] == pytest.approx(1.0, abs=1e-5)


def test_population_update_drops_fingerprint_duplicates():
    cfg = ev.GAConfig(counts=(16, 8))
    rng = np.random.default_rng(8)
    a = ev.random_structure(cfg, rng)
    b = ev.random_structure(cfg, rng)
    a_copy = (a[0], a[1], a[2][np.concatenate([rng.permutation(16), 16 + rng.permutation(8)])] + 0.5)
    structs = [a, b, a_copy]
    E = np.array([-1.50, -1.49, -1.5005])    # the copy within the energy window of its original
    fp = ev.fingerprints(structs, 2, cfg)
    keep = ev.unique_by_fingerprint(E, fp, ev.cell_volumes(structs), cfg.dup_cut, 10)
    assert sorted(keep) == [1, 2]             # the lower-energy copy stays, its duplicate goes, b stays
    # with the test switched off every finite structure stays; infinite energies always go
    E2 = np.array([-1.50, np.inf, -1.5005])
    assert sorted(ev.unique_by_fingerprint(E2, fp, ev.cell_volumes(structs), 1.0, 10)) == [0, 2]


def test_duplicate_walk_shortcut_agrees_with_the_full_walk():
    """Skipping the comparisons among a block already free of duplicates (the previous population) changes nothing."""
    cfg = ev.GAConfig(counts=(16, 8))
    rng = np.random.default_rng(12)
    old = [ev.random_structure(cfg, rng) for _ in range(40)]
    new = [ev.random_structure(cfg, rng) for _ in range(20)]
    # plant duplicates of old members among the new ones (translated copies, energies within the window)
    for q in (3, 7, 11):
        s = old[q]
        new.append((s[0], s[1], s[2] + 0.3))
    structs = old + new
    E = np.concatenate([rng.normal(-1.7, 0.02, 40), rng.normal(-1.7, 0.02, 20), [0, 0, 0]])
    E[60:] = E[[3, 7, 11]] + 0.001
    fp = ev.fingerprints(structs, 2, cfg)
    V = ev.cell_volumes(structs)
    full = ev.unique_by_fingerprint(E, fp, V, 0.95, 50)
    fast = ev.unique_by_fingerprint(E, fp, V, 0.95, 50, n_unique=40)
    assert full == fast
    assert not (set(full) & {60, 61, 62})


def _walk_reference(E, fp, V, cut, limit, n_unique=0):
    """The duplicate walk written out pair by pair (every kept structure scored with COMP_CL, no windows)."""
    order = np.argsort(E, kind="stable")
    keep = []
    for k in order:
        if not np.isfinite(E[k]):
            continue
        dup = False
        for j in keep:
            if k < n_unique and j < n_unique:
                continue
            w = ev.comp_cl(fp[k], fp[j][None], E[k], np.array([E[j]]), V[k], np.array([V[j]]))[0]
            if w > cut:
                dup = True
                break
        if not dup:
            keep.append(int(k))
        if len(keep) == limit:
            break
    return keep


@pytest.mark.parametrize("n_unique", [0, 30])
def test_native_duplicate_walk_equals_pairwise_comp_cl(n_unique):
    """mg_evo_unique (energy/volume windows, native walk) against COMP_CL evaluated for every kept pair, on
    structures with planted duplicates, tied energies and an infinite energy."""
    cfg = ev.GAConfig(counts=(16, 8))
    rng = np.random.default_rng(21)
    base = [ev.random_structure(cfg, rng) for _ in range(45)]
    dups = [(s[0], s[1], s[2] + rng.normal(0.0, 0.002, s[2].shape)) for s in base[:8]]
    structs = base + dups
    E = np.concatenate([rng.normal(-1.7, 0.01, 45), np.zeros(8)])
    E[45:] = E[:8] + rng.uniform(-0.002, 0.002, 8)
    E[10] = E[11]            # a tie
    E[20] = np.inf           # dropped
    fp = ev.fingerprints(structs, 2, cfg)
    V = ev.cell_volumes(structs)
    got = ev.unique_by_fingerprint(E, fp, V, 0.95, 40, n_unique=n_unique)
    want = _walk_reference(E, fp, V, 0.95, 40, n_unique=n_unique)
    assert got == want
    assert 20 not in got and len(set(range(45, 53)) & set(got)) + len(set(range(8)) & set(got)) <= 8
