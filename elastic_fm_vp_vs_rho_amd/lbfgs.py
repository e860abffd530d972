"""L-BFGS history and two-loop recursion with the interface of the reference's
``lbfgs.py`` (``compute_sk_yk`` L:12-25, ``L_BFGS`` L:27-67).

History layout (kept exactly, it is part of the interface): ``sk``/``yk`` are
``(L, ny, nz, nx)`` arrays.  Slot 0 always holds the current model / gradient;
slots 1 .. L-2 hold the most recent (Δm, Δg) pairs, newest first.  The shift
that makes room for a new pair stops at slot L-2, so slot L-1 is never written
and the effective history is L-2 pairs.

The recursion is the standard two-loop form (Nocedal & Wright, Alg. 7.4) over
``layers = min(iter, L-1)`` slots with the initial scaling
γ = s₁·y₁ / y₁·y₁; zero denominators make the corresponding factor 0.
Inner products are accumulated in float64.
"""
from __future__ import annotations

import numpy as np


def _dot(a, b):
    return float(np.dot(np.ravel(a).astype(np.float64, copy=False), np.ravel(b).astype(np.float64, copy=False)))


def compute_sk_yk(sk, yk, mk, gk, lbfgsLayers, iter):
    """Update the history in place with the model ``mk`` and gradient ``gk`` of
    iteration ``iter``."""
    L = lbfgsLayers
    if iter == 0:
        sk[0] = mk
        yk[0] = gk
        return
    for hist, cur in ((sk, mk), (yk, gk)):
        newest = cur - hist[0]                 # Δ against the previous current value
        if L > 2:                              # pairs live in slots 1 .. L-2
            hist[2:L - 1] = hist[1:L - 2].copy()
            hist[1] = newest
        hist[0] = cur


def L_BFGS(sk, yk, gk, lbfgsLayers, iter):
    """Two-loop recursion: returns r ≈ H·gk for the stored history."""
    L = lbfgsLayers
    layers = iter if iter < L else L - 1
    yy = _dot(yk[1], yk[1]) if L > 1 else 0.0
    gamma = 0.0 if yy == 0 else _dot(sk[1], yk[1]) / yy
    q = np.array(gk, copy=True)
    sigma = np.zeros(L)
    eps = np.zeros(L)
    for i in range(layers):
        ys = _dot(yk[i + 1], sk[i + 1])
        sigma[i] = 0.0 if ys == 0 else 1.0 / ys
        eps[i] = sigma[i] * _dot(sk[i + 1], q)
        q = q - eps[i] * yk[i + 1]
    r = gamma * q
    for i in range(layers, 0, -1):
        beta = sigma[i - 1] * _dot(yk[i], r)
        r = r + sk[i] * (eps[i - 1] - beta)
    return r
