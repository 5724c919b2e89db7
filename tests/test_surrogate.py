"""CPU tests of the screening surrogate's oracle and of the evolutionary operators.

The FP64 surrogate oracle (oracle/surrogate_oracle.py) is pinned here: its gradient against central finite
differences of its loss, its Adam step against the closed form of the first step, and its features against a
per-species mean computed another way.  The GA operators are checked for the invariants the reference's
BULK_MATE keeps (atom count, composition, species grouping, one parent's lattice, the 10% volume window) and
the periodic minimum distance against a brute-force search over a larger image range.
"""
import numpy as np
import pytest

from oracle import surrogate_oracle as so


def _rand_params(nin, rng, scale=0.5):
    return rng.normal(0.0, scale, so.nparams(nin))


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    nin, B = 7, 5
    p = _rand_params(nin, rng)
    X = rng.normal(size=(B, nin))
    y = rng.normal(size=B)
    mu, sig = rng.normal(size=nin) * 0.1, 0.5 + rng.random(nin)
    ymu, ysig = 0.3, 1.7
    L, g = so.loss_grad(p, nin, X, y, mu, sig, ymu, ysig)
    h = 1e-6
    # every bias and output weight, and a sample of the weight matrices
    idx = list(range(so.H * nin, so.H * nin + so.H)) + list(range(len(p) - 2 * so.H - 1, len(p)))
    idx += list(rng.choice(so.H * nin, 40, replace=False)) + list(so.H * nin + so.H + rng.choice(so.H * so.H, 40, replace=False))
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (so.loss_grad(pp, nin, X, y, mu, sig, ymu, ysig)[0] - so.loss_grad(pm, nin, X, y, mu, sig, ymu, ysig)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-6, abs=1e-9), k


def test_oracle_forward_is_the_textbook_network():
    rng = np.random.default_rng(1)
    nin = 5
    p = _rand_params(nin, rng)
    x = rng.normal(size=nin)
    W0, b0, W1, b1, w2, b2 = so.unpack(p, nin)
    # written out unit by unit
    h1 = [np.tanh(sum(W0[m, n] * x[n] for n in range(nin)) + b0[m]) for m in range(so.H)]
    h2 = [np.tanh(sum(W1[m, n] * h1[n] for n in range(so.H)) + b1[m]) for m in range(so.H)]
    ref = sum(w2[m] * h2[m] for m in range(so.H)) + b2
    y, _ = so.forward(p, nin, x[None, :], np.zeros(nin), np.ones(nin))
    assert y[0] == pytest.approx(ref, rel=1e-12)
    # normalisation: y = ymu + ysig * f((x - mu)/sig)
    mu, sig = rng.normal(size=nin), 1 + rng.random(nin)
    y2, _ = so.forward(p, nin, (x * sig + mu)[None, :], mu, sig, 2.0, 3.0)
    assert y2[0] == pytest.approx(2.0 + 3.0 * ref, rel=1e-12)


def test_oracle_adam_first_step_closed_form():
    rng = np.random.default_rng(2)
    p = rng.normal(size=50)
    g = rng.normal(size=50)
    lr, eps = 1e-3, 1e-8
    p1, m, v = so.adam_step(p, g, np.zeros(50), np.zeros(50), 1, lr, eps=eps)
    assert np.allclose(p1, p - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)
    # second step with a constant gradient: bias-corrected moments equal g and g^2 again
    p2, _, _ = so.adam_step(p1, g, m, v, 2, lr, eps=eps)
    assert np.allclose(p2, p1 - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)


def test_oracle_features_are_species_means():
    rng = np.random.default_rng(3)
    D = 20
    atom_off = [0, 5, 9, 12]
    spc = np.array([0, 0, 1, 1, 1, 0, 0, 0, 0, 1, 1, 1])   # the third structure has no atom of species 0
    G = rng.normal(size=(12, D))
    f = so.features(G, spc, atom_off, 2)
    assert f.shape == (3, 2 * D)
    for s in range(3):
        a, b = atom_off[s], atom_off[s + 1]
        for q in range(2):
            sel = spc[a:b] == q
            exp = G[a:b][sel].sum(axis=0) / max(sel.sum(), 1)
            assert np.allclose(f[s, q * D:(q + 1) * D], exp if sel.any() else 0.0)
    assert not f[2, :D].any()


# ------------------------------------------------------------------ evolutionary operators
def _parents(rng):
    from maise_amd.evolve import GAConfig, random_structure
    cfg = GAConfig(counts=(6, 3), vol_range=(24.0, 30.0), sigma=0.1)
    return cfg, [random_structure(cfg, rng) for _ in range(6)]


def _lat_match(lat, parents):
    from maise_amd.evolve import lattice_params
    lc, ac = lattice_params(lat)
    for _, L, _ in parents:
        l2, a2 = lattice_params(L)
        for k in range(3):
            perm = [(q + k) % 3 for q in range(3)]
            if np.allclose(lc, l2[perm]) and np.allclose(ac, a2[perm]):
                return L
    return None


def test_offspring_keep_count_composition_lattice_and_dmin():
    """Native crossover + mutation (mg_evo_offspring): every child has the parents' atom count and composition,
    species grouped, the lattice of one parent (cyclically permuted), a volume within 10% of the mean of that
    parent and another, and all periodic distances >= dmin."""
    from maise_amd.evolve import make_offspring, min_distance
    rng = np.random.default_rng(4)
    cfg, par = _parents(rng)
    kids = make_offspring(par, np.ones(len(par)), cfg, np.random.default_rng(7), 64)
    vols = [abs(np.linalg.det(p[1])) for p in par]
    for spc, lat, x in kids:
        assert len(spc) == len(par[0][0])
        assert np.array_equal(np.bincount(spc), np.bincount(par[0][0]))
        assert np.all(np.diff(spc) >= 0)
        assert _lat_match(lat, par) is not None
        V = abs(np.linalg.det(lat))
        assert any(abs(V - 0.5 * (V + v)) / (0.5 * (V + v)) < 0.1 for v in vols)
        assert min_distance(lat, x) >= cfg.dmin


def test_min_distance_brute_force():
    from maise_amd.evolve import min_distance
    from maise_amd.workloads import random_triclinic
    rng = np.random.default_rng(5)
    for _ in range(5):
        _, lat, x = random_triclinic(7, 1, rng, dmin=1.5)
        best = np.inf
        rng_img = range(-3, 4)
        for i in range(len(x)):
            for j in range(len(x)):
                for a in rng_img:
                    for b in rng_img:
                        for c in rng_img:
                            if i == j:
                                continue
                            d = x[j] - x[i] + np.array([a, b, c]) @ lat
                            best = min(best, float(np.linalg.norm(d)))
        assert min_distance(lat, x) == pytest.approx(best, rel=1e-12)
        assert best >= 1.5 - 1e-9


def test_offspring_respect_dmin_and_are_seeded():
    from maise_amd.evolve import make_offspring, min_distance
    rng = np.random.default_rng(6)
    cfg, par = _parents(rng)
    w = np.ones(len(par))
    a = make_offspring(par, w, cfg, np.random.default_rng(9), 8)
    b = make_offspring(par, w, cfg, np.random.default_rng(9), 8)
    for (s1, l1, x1), (s2, l2, x2) in zip(a, b):
        assert np.array_equal(x1, x2) and np.array_equal(l1, l2)
        assert min_distance(l1, x1) >= cfg.dmin
