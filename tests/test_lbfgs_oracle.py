"""The L-BFGS restatement in the oracle (oracle/pyoracle.py), checked on the CPU before the GPU relaxation is
checked against it: the two-loop recursion against the inverse-Hessian BFGS update written out as matrices
(Nocedal & Wright, Numerical Optimization, eqs. 6.17 and 7.19), and a relaxation on FP64 oracle forces that
must end below the force threshold."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def _explicit_h(S, Y, gamma):
    """H_k from H_0 = gamma I by H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T, oldest pair first."""
    n = len(S[0])
    H = gamma * np.eye(n)
    for s, y in zip(S, Y):
        rho = 1.0 / (y @ s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return H


@pytest.mark.parametrize("npairs", [0, 1, 3, 7])
def test_two_loop_equals_explicit_bfgs_update(npairs):
    from oracle.pyoracle import lbfgs_direction
    rng = np.random.default_rng(11 + npairs)
    n = 9
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)             # SPD Hessian: y = A s has y.s > 0
    S = [rng.standard_normal(n) for _ in range(npairs)]
    Y = [A @ s for s in S]
    gamma = (S[-1] @ Y[-1]) / (Y[-1] @ Y[-1]) if npairs else 0.02
    q = rng.standard_normal(n)
    got = lbfgs_direction(S, Y, gamma, q)
    want = _explicit_h(S, Y, gamma) @ q if npairs else gamma * q
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    if npairs:
        # secant condition of the newest pair: H y = s
        assert np.allclose(lbfgs_direction(S, Y, gamma, Y[-1]), S[-1], rtol=1e-10, atol=1e-12)


def test_two_loop_with_n_conjugate_pairs_inverts_a_quadratic():
    """n pairs along A-conjugate directions of a quadratic give H = A^-1 whatever gamma (BFGS's quadratic
    termination property), so H g recovers the Newton step exactly."""
    from oracle.pyoracle import lbfgs_direction
    rng = np.random.default_rng(5)
    n = 6
    B = rng.standard_normal((n, n))
    A = B @ B.T + np.eye(n)
    # A-conjugate directions by Gram-Schmidt in the A inner product
    S = []
    for v in rng.standard_normal((n, n)):
        for s in S:
            v = v - (v @ A @ s) / (s @ A @ s) * s
        S.append(v)
    Y = [A @ s for s in S]
    q = rng.standard_normal(n)
    assert np.allclose(lbfgs_direction(S, Y, 0.37, q), np.linalg.solve(A, q), rtol=1e-9, atol=1e-10)


def test_lbfgs_restatement_relaxes_a_cell():
    from maise_amd import workloads as wl
    from oracle import pyoracle as po
    m = po.OracleModel(os.path.join(GOLDEN, "models", "nn_CuPdAg_d3_v1"))
    rng = np.random.default_rng(31)
    spc, lat, x = wl.random_triclinic(12, 3, rng)
    X, E, steps, ok = po.lbfgs_relax(m, spc, lat, x, fmax=0.01, max_steps=600, urep=100.0)
    assert ok
    o = m.evaluate(spc, lat, X, urep=100.0)
    assert np.abs(o["F"]).max() < 0.0101
    e0 = m.evaluate(spc, lat, x, urep=100.0)["E"]
    assert E < e0
