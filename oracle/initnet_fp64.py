"""FP64 NumPy restatement of the learned MT/location initializer network
(test infrastructure: only tests/ may import it, as the checker of the HIP
training and inference kernels).

The network has no reference counterpart.  Architecture: ``nblocks`` strided
1-D convolutions (``channels`` filters of width ``ks``, stride, padding) each
followed by bias and ReLU, a flatten, a dense hidden layer with ReLU and a dense
output layer; loss = mean squared error.  Activations are (C, B, L).

With ``emulate=True`` the restatement rounds to BF16 exactly where the device
path does (weights, the network input, every post-ReLU activation that feeds a
GEMM, and the gradient operands of the backward GEMMs), and accumulates in
FP64; the device accumulates in FP32, so the two agree to FP32 accumulation
accuracy plus the rare BF16 rounding-boundary flip.  With ``emulate=False`` it
is the plain FP64 network, which the finite-difference tests pin.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def bf16(x):
    """Round to BF16 (nearest even, via float32) and return float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    out = (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    out[np.isnan(a)] = np.nan
    return out


@dataclass(frozen=True)
class Spec:
    cin: int = 192
    length: int = 1024
    channels: int = 64
    ks: int = 9
    stride: int = 4
    pad: int = 4
    nblocks: int = 4
    hidden: int = 128
    nout: int = 9

    def convs(self):
        """[(cin, cout, lin, lout)] per convolution block."""
        out, c, L = [], self.cin, self.length
        for _ in range(self.nblocks):
            lo = (L + 2 * self.pad - self.ks) // self.stride + 1
            out.append((c, self.channels, L, lo))
            c, L = self.channels, lo
        return out

    @property
    def flat(self):
        return self.channels * self.convs()[-1][3]

    def shapes(self):
        sh = {}
        for i, (ci, co, _, _) in enumerate(self.convs()):
            sh[f"W{i}"] = (co, ci * self.ks)
            sh[f"b{i}"] = (co,)
        sh["F1"] = (self.hidden, self.flat)
        sh["c1"] = (self.hidden,)
        sh["F2"] = (self.nout, self.hidden)
        sh["c2"] = (self.nout,)
        return sh


def init_params(spec: Spec, seed=0):
    """He-normal weights (the output layer scaled by 1/sqrt(fan_in)), zero biases, float32."""
    rng = np.random.default_rng(seed)
    p = {}
    for k, shp in spec.shapes().items():
        if k[0] in "bc":
            p[k] = np.zeros(shp, np.float32)
        else:
            gain = 1.0 if k == "F2" else 2.0
            p[k] = (rng.standard_normal(shp) * np.sqrt(gain / shp[1])).astype(np.float32)
    return p


def im2col(x, ks, stride, pad):
    C, B, L = x.shape
    lo = (L + 2 * pad - ks) // stride + 1
    xp = np.zeros((C, B, L + 2 * pad + stride), x.dtype)
    xp[:, :, pad:pad + L] = x
    col = np.empty((C, ks, B, lo), x.dtype)
    for k in range(ks):
        col[:, k] = xp[:, :, k:k + stride * lo:stride]
    return col.reshape(C * ks, B * lo)


def col2im(dcol, C, B, L, ks, stride, pad):
    lo = (L + 2 * pad - ks) // stride + 1
    d = dcol.reshape(C, ks, B, lo)
    xp = np.zeros((C, B, L + 2 * pad + stride), dcol.dtype)
    for k in range(ks):
        xp[:, :, k:k + stride * lo:stride] += d[:, k]
    return xp[:, :, pad:pad + L]


def forward(spec: Spec, p, x, emulate=True):
    """x (cin, B, length) -> out (nout, B), cache."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    B = x.shape[1]
    h = r(x)
    cache = {"x0": h}
    for i, (ci, co, L, lo) in enumerate(spec.convs()):
        col = im2col(h, spec.ks, spec.stride, spec.pad)
        y = np.maximum(r(p[f"W{i}"]) @ col + np.asarray(p[f"b{i}"], np.float64)[:, None], 0.0)
        cache[f"col{i}"], cache[f"y{i}"] = col, y
        h = r(y).reshape(co, B, lo)
    X = h.transpose(0, 2, 1).reshape(spec.flat, B)
    z = np.maximum(r(p["F1"]) @ X + np.asarray(p["c1"], np.float64)[:, None], 0.0)
    zb = r(z)
    out = r(p["F2"]) @ zb + np.asarray(p["c2"], np.float64)[:, None]
    cache.update(X=X, z=z, zb=zb, out=out)
    return out, cache


def backward(spec: Spec, p, cache, target, emulate=True):
    """Mean-squared-error loss and gradients of every parameter."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    out = cache["out"]
    n = out.size
    res = out - target
    loss = float(np.sum(res * res) / n)
    g = {}
    dout = 2.0 * res / n
    db = r(dout)
    g["F2"] = db @ cache["zb"].T
    g["c2"] = dout.sum(1)
    dz = r(p["F2"]).T @ db
    dpre = dz * (cache["z"] > 0)
    dpb = r(dpre)
    g["F1"] = dpb @ cache["X"].T
    g["c1"] = dpre.sum(1)
    dX = r(p["F1"]).T @ dpb
    convs = spec.convs()
    B = out.shape[1]
    co, lo = convs[-1][1], convs[-1][3]
    dy = dX.reshape(co, lo, B).transpose(0, 2, 1).reshape(co, B * lo)
    for i in range(len(convs) - 1, -1, -1):
        ci, co, L, lo = convs[i]
        dpre = dy * (cache[f"y{i}"] > 0)
        dpb = r(dpre)
        g[f"W{i}"] = dpb @ cache[f"col{i}"].T
        g[f"b{i}"] = dpre.sum(1)
        if i > 0:
            dcol = r(p[f"W{i}"]).T @ dpb
            dy = col2im(dcol, ci, B, L, spec.ks, spec.stride, spec.pad).reshape(ci, B * L)
    return loss, g


def adam(w, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (float64), returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1.0 / (1 - b1 ** t), 1.0 / (1 - b2 ** t)
    return w - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v


def pack_gathers(g, L):
    """(S, ncomp, ntmax, nrec) -> (S, ncomp*nrec, L) normalised by each shot's max |g|, BF16."""
    S, nc, nt, nr = g.shape
    mx = np.abs(g.reshape(S, -1)).max(1)
    inv = np.where(mx > 0, 1.0 / np.where(mx > 0, mx, 1.0), 0.0).astype(np.float32)
    x = (g[:, :, :L, :].astype(np.float32) * inv[:, None, None, None]).transpose(0, 1, 3, 2)
    return bf16(x.reshape(S, nc * nr, L))
