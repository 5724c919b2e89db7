"""GPU tests of the screening surrogate: features, gradients, Adam, prediction and a short screened search.

Every kernel result is compared with the FP64 surrogate oracle on the same inputs (FP32 tolerances stated per
test).  The short search checks the plumbing of the screened generation loop and that the trained surrogate
ranks offspring better than chance.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

MGCA = os.path.join(GOLDEN, "models", "nn_MgCa_d3_v0")


def _sur(nin, seed=3):
    from maise_amd.surrogate import Surrogate
    return Surrogate(nin, seed=seed)


def test_features_match_oracle_means():
    import maise_amd
    from maise_amd.evolve import GAConfig, random_structure
    from oracle import pyoracle as po
    from oracle import surrogate_oracle as so
    print("[gpu case] surrogate features, 6 Mg-Ca structures", flush=True)
    rng = np.random.default_rng(11)
    cfg = GAConfig(counts=(16, 8))
    structs = [random_structure(cfg, rng) for _ in range(6)]
    m = maise_amd.Model(MGCA)
    c = maise_amd.Context(m, urep=100.0)
    c.set_batch(structs)
    c.eval(forces=False)
    f = c.features()
    assert f.shape == (6, 2 * m.D)
    G = c.descriptors()
    spc = np.concatenate([s[0] for s in structs])
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in structs])])
    ref32 = so.features(G, spc, off, m.nspc)
    assert np.allclose(f, ref32, rtol=1e-5, atol=1e-7)
    # and against FP64 oracle descriptors
    om = po.OracleModel(MGCA)
    G64 = np.concatenate([om.evaluate(*s, forces=False, want_g=True)["G"] for s in structs])
    ref64 = so.features(G64, spc, off, m.nspc)
    assert np.allclose(f, ref64, rtol=1e-4, atol=1e-6)
    c.close()


def test_gradient_and_prediction_match_oracle():
    from oracle import surrogate_oracle as so
    print("[gpu case] surrogate gradient nin=102 B=333", flush=True)
    rng = np.random.default_rng(12)
    nin, B = 102, 333
    s = _sur(nin)
    X = rng.normal(size=(B, nin)).astype(np.float32)
    y = rng.normal(size=B).astype(np.float32)
    mu = (0.1 * rng.normal(size=nin)).astype(np.float32)
    sig = (0.5 + rng.random(nin)).astype(np.float32)
    s.set_norm(mu, sig, 0.25, 1.5)
    p = s.params.astype(np.float64)
    loss, g = s.grad(X, y)
    L64, g64 = so.loss_grad(p, nin, X.astype(np.float64), y.astype(np.float64), mu.astype(np.float64),
                            sig.astype(np.float64), 0.25, 1.5)
    assert loss == pytest.approx(L64, rel=1e-5)
    scale = np.abs(g64).max()
    assert np.abs(g - g64).max() <= 1e-5 * scale + 1e-7
    pred = s.predict(X)
    y64, _ = so.forward(p, nin, X.astype(np.float64), mu.astype(np.float64), sig.astype(np.float64), 0.25, 1.5)
    assert np.abs(pred - y64).max() < 1e-5 * max(1.0, np.abs(y64).max())
    s.close()


def test_adam_steps_match_oracle():
    from oracle import surrogate_oracle as so
    print("[gpu case] surrogate Adam, 3 full-batch steps", flush=True)
    rng = np.random.default_rng(13)
    nin, B = 40, 96
    s = _sur(nin, seed=5)
    X = rng.normal(size=(B, nin)).astype(np.float32)
    y = (X[:, 0] - 0.5 * X[:, 1] ** 2).astype(np.float32)
    p = s.params.astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    lr = 1e-3
    for t in range(1, 4):
        # one epoch with one batch of all samples = one Adam step on the full-batch gradient
        s.train(X, y, epochs=1, batch=B, lr=lr, seed=t)
        _, g = so.loss_grad(p, nin, X.astype(np.float64), y.astype(np.float64), np.zeros(nin), np.ones(nin))
        p, m, v = so.adam_step(p, g, m, v, t, lr)
        # the update is ~lr per step; FP32 gradients near zero may differ in sign, moving those few weights
        # by at most 2 lr per step
        d = np.abs(s.params - p)
        assert np.median(d) < 1e-6
        assert d.max() <= 2 * lr * t + 1e-6
    s.close()


def test_training_fits_a_smooth_function():
    print("[gpu case] surrogate training on a smooth target", flush=True)
    rng = np.random.default_rng(14)
    nin, n = 16, 2000
    s = _sur(nin, seed=7)
    X = rng.normal(size=(n, nin)).astype(np.float32)
    y = (np.sin(X[:, 0]) + 0.3 * X[:, 1] * X[:, 2] - 0.2 * X[:, 3]).astype(np.float32)
    s.fit_norm(X, y)
    l0, _ = s.grad(X, y)
    l1 = s.train(X, y, epochs=30, batch=64, lr=2e-3, seed=1)
    assert l1 < 0.05 * l0
    Xt = rng.normal(size=(500, nin)).astype(np.float32)
    yt = np.sin(Xt[:, 0]) + 0.3 * Xt[:, 1] * Xt[:, 2] - 0.2 * Xt[:, 3]
    assert np.sqrt(np.mean((s.predict(Xt) - yt) ** 2)) < 0.3 * yt.std()
    s.close()


def test_training_left_in_flight_matches_blocking_training():
    """mg_sur_train_begin/_end (the run enqueued, the caller's arrays released at once, host work in between, a
    prediction that waits for the run) gives the parameters and loss of mg_sur_train bit for bit."""
    print("[gpu case] surrogate training begin/end vs blocking", flush=True)
    rng = np.random.default_rng(15)
    nin, n = 24, 700
    X = rng.normal(size=(n, nin)).astype(np.float32)
    y = (np.cos(X[:, 0]) + 0.2 * X[:, 1]).astype(np.float32)
    a, b = _sur(nin, seed=9), _sur(nin, seed=9)
    for s in (a, b):
        s.fit_norm(X, y)
    la = a.train(X, y, epochs=7, batch=32, lr=1e-3, seed=4)
    Xc, yc = X.copy(), y.copy()
    b.train_begin(Xc, yc, epochs=7, batch=32, lr=1e-3, seed=4)
    Xc[:] = 0.0                                    # the caller's buffers are free once begin returns
    yc[:] = 0.0
    pb = b.predict(X[:10])                         # waits for the run in flight
    lb = b.train_end()
    assert lb == la
    assert np.array_equal(a.params, b.params)
    assert np.array_equal(a.predict(X[:10]), pb)
    a.close()
    b.close()


def test_screened_search_short():
    from maise_amd.evolve import GAConfig, SurrogateGA
    print("[gpu case] screened GA, Mg16Ca8, pop 96, 3 generations", flush=True)
    cfg = GAConfig(pop=96, generations=3, seed=21, max_steps=1500)
    ga = SurrogateGA(MGCA, cfg)
    res = ga.run()
    ga.close()
    reps = res.reports
    assert len(reps) == 4
    assert reps[1].relaxed == round(0.25 * 96) + round(0.10 * 96)
    assert all(np.isfinite(r.best_eV_atom) for r in reps)
    assert all(reps[k + 1].best_eV_atom <= reps[k].best_eV_atom + 1e-12 for k in range(3))  # elitist
    rec = np.mean([r.recall_top25 for r in reps[1:]])
    # a selection of 35% of the offspring drawn at random recalls 35% of the top quarter on average
    print("recall@25%", [r.recall_top25 for r in reps[1:]], "MAE", [r.mae_eV_atom for r in reps[1:]], flush=True)
    assert rec > 0.35


def test_screened_search_ranks_prerelaxed_offspring():
    """prerelax: every offspring takes a few FIRE steps (one batch) before the surrogate ranks it, and the kept ones
    go on from there.  The kept share is unchanged, the search stays elitist, and the audit (every offspring relaxed
    from scratch) scores the ranking; on the 512-population workload 40 steps raised recall@25% from 0.39 to 0.53
    (measured in the bench's search leg), here the recall is printed."""
    from maise_amd.evolve import GAConfig, SurrogateGA
    print("[gpu case] screened GA with prerelax 20, Mg16Ca8, pop 96, 3 generations", flush=True)
    cfg = GAConfig(pop=96, generations=3, seed=21, max_steps=1500, prerelax=20)
    ga = SurrogateGA(MGCA, cfg)
    res = ga.run()
    ga.close()
    reps = res.reports
    assert len(reps) == 4
    assert reps[1].relaxed == round(0.25 * 96) + round(0.10 * 96)
    assert all(r.seconds_prerelax > 0 for r in reps[1:])
    assert all(np.isfinite(r.best_eV_atom) for r in reps)
    assert all(reps[k + 1].best_eV_atom <= reps[k].best_eV_atom + 1e-12 for k in range(3))
    print("recall@25% after 20 steps", [r.recall_top25 for r in reps[1:]], flush=True)
    assert all(0.0 <= r.recall_top25 <= 1.0 for r in reps[1:])


def test_capped_relaxation_carries_the_tail_over():
    """With a FIRE cap per generation, unconverged structures move into the next generation's batch instead of
    holding every generation to the slowest structure; nothing is lost except past max_steps in total."""
    from maise_amd.evolve import GAConfig, SurrogateGA
    print("[gpu case] GA with relax_cap 60, Mg16Ca8, pop 64, 4 generations", flush=True)
    cfg = GAConfig(pop=64, generations=4, seed=23, relax_cap=60, max_steps=400, audit=False)
    ga = SurrogateGA(MGCA, cfg)
    res = ga.run()
    ga.close()
    reps = res.reports
    assert any(r.carried > 0 for r in reps)                  # 60 steps leave some structures unconverged
    assert all(np.isfinite(r.best_eV_atom) for r in reps)
    assert all(reps[k + 1].best_eV_atom <= reps[k].best_eV_atom + 1e-12 for k in range(4))
    assert np.isfinite(res.population_E).all()


def test_screening_at_equal_search_time():
    """The same seeded search (population 512, as the search workload) with and without the surrogate for the
    same search clock.  What screening delivers by construction is asserted: 35% of the offspring relaxed.  The
    generation rate is printed, not asserted: a batch of 179 structures of 24 atoms relaxes in about 60% of the
    time of 512, and the surrogate's evaluation and training take about what that saves, so at this population the
    two searches run about as many generations (the bench's comparison at population 2,048, where relaxation time
    grows with the batch, gives the screened search 1.6x the generations).  Which search reaches the lower energy
    first is measured and printed, not asserted: it depends on the seed; the audited run above asserts a recall
    above random selection's."""
    from maise_amd.evolve import GAConfig, compare_screening
    print("[gpu case] equal-time search, Mg16Ca8, pop 512, 1 seed x 4 s", flush=True)
    cfg = GAConfig(pop=512, relax_cap=300)
    out = compare_screening(MGCA, cfg, 4.0, [31],
                            variants={"unscreened": {"screen": False}, "screened": {"screen": True}})
    for name, s in out["summary"].items():
        print(name, s, flush=True)
    for name, rr in out["runs"].items():
        for r in rr:
            print(name, r["seed"], "generations", r["generations"], "best", round(r["best_eV_atom"], 5),
                  "seconds to target", r["seconds_to_target"], flush=True)
    for r in out["runs"]["screened"]:
        assert r["relaxed_per_generation"] == round(0.25 * 512) + round(0.10 * 512)
    for name in ("unscreened", "screened"):
        for r in out["runs"][name]:
            assert r["generations"] >= 3 and np.isfinite(r["best_eV_atom"])
