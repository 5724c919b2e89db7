"""FP64 restatement of the screening surrogate -- TEST INFRASTRUCTURE ONLY.

Only tests/ (and bench.py's CPU-baseline leg) may import this module; the product path trains and evaluates
the surrogate with its own HIP kernels and never calls into here.

The network is nin-64-64-1 tanh with a linear output, parameters flat in the order
W0[64][nin] b0[64] W1[64][64] b1[64] w2[64] b2 (per-layer B then W blocks as the reference stores a network,
src/nmlp.c:591-608), inputs standardised x' = (x - mu)/sig, target t' = (t - ymu)/ysig, loss
L = (1/B) sum_b (y'_b - t'_b)^2.  The gradient is the reference's training chain DE_ANN_PARA
(src/nmlp.c:292-354: output bias/weights from the loss derivative times the last activations, then the
derivative pushed through W and tanh' = 1 - e^2 layer by layer), restated for this loss; Adam is Kingma & Ba
(2015), Algorithm 1.

Pinned by tests/test_surrogate.py: the gradient against central finite differences of the loss, the first Adam
step against its closed form lr*g/(|g|+eps), and the features against a per-species mean written out with
explicit loops.
"""
from __future__ import annotations

import numpy as np

H = 64


def nparams(nin: int) -> int:
    return H * nin + H * H + 3 * H + 1


def unpack(p: np.ndarray, nin: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W0 = p[o:o + H * nin].reshape(H, nin); o += H * nin
    b0 = p[o:o + H]; o += H
    W1 = p[o:o + H * H].reshape(H, H); o += H * H
    b1 = p[o:o + H]; o += H
    w2 = p[o:o + H]; o += H
    b2 = p[o]
    return W0, b0, W1, b1, w2, b2


def forward(p, nin, X, mu, sig, ymu=0.0, ysig=1.0):
    """Predictions in target units and the activations (x', h1, h2, y')."""
    W0, b0, W1, b1, w2, b2 = unpack(p, nin)
    xn = (np.asarray(X, dtype=np.float64) - mu) / sig
    h1 = np.tanh(xn @ W0.T + b0)
    h2 = np.tanh(h1 @ W1.T + b1)
    yn = h2 @ w2 + b2
    return ymu + ysig * yn, (xn, h1, h2, yn)


def loss_grad(p, nin, X, y, mu, sig, ymu=0.0, ysig=1.0):
    """Mean squared error in normalised units and its gradient with respect to the flat parameters."""
    W0, b0, W1, b1, w2, b2 = unpack(p, nin)
    _, (xn, h1, h2, yn) = forward(p, nin, X, mu, sig, ymu, ysig)
    t = (np.asarray(y, dtype=np.float64) - ymu) / ysig
    B = len(t)
    diff = yn - t
    L = float(np.mean(diff * diff))
    e = 2.0 * diff / B                        # dL/dy'
    g_b2 = e.sum()
    g_w2 = e @ h2
    d2 = (e[:, None] * w2[None, :]) * (1.0 - h2 * h2)
    g_W1 = d2.T @ h1
    g_b1 = d2.sum(axis=0)
    d1 = (d2 @ W1) * (1.0 - h1 * h1)
    g_W0 = d1.T @ xn
    g_b0 = d1.sum(axis=0)
    g = np.concatenate([g_W0.ravel(), g_b0, g_W1.ravel(), g_b1, g_w2, [g_b2]])
    return L, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step t >= 1); returns new (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def features(G, spc, atom_off, nspc):
    """Per structure and species, the mean of the whole descriptor row over that species' atoms; zeros for an
    absent species."""
    G = np.asarray(G, dtype=np.float64)
    D = G.shape[1]
    S = len(atom_off) - 1
    out = np.zeros((S, nspc * D))
    for s in range(S):
        for q in range(nspc):
            rows = [i for i in range(atom_off[s], atom_off[s + 1]) if spc[i] == q]
            if rows:
                out[s, q * D:(q + 1) * D] = G[rows].mean(axis=0)
    return out
