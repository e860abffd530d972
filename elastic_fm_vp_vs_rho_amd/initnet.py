"""Learned moment-tensor / location initializer: a 1-D convolutional network
trained and run on the GPU by the library's own kernels (include/efm_initnet.h).

No counterpart in the reference, whose inversions start from the fixed moment
tensor Mmod = (0.5, 0, 0.5, 0, 0.5, 0) (mainPropElastic3D_MT.py M:2777) at the
catalogue location; the network proposes both from the recorded gathers and the
plain start remains the fallback.

Input: gathers of ``nrec`` receivers × 3 components × ``length`` samples, each
shot normalised by its peak amplitude, as ``cin = 3·nrec`` channels (BF16).
Network: ``nblocks`` × [conv1d(ks, stride, pad) → bias → ReLU] with
``channels`` filters, flatten, dense(hidden) → ReLU, dense(nout).  Outputs: the
6 components of the Frobenius-normalised moment tensor and the source x, y, z
scaled to [-1, 1] over the sampled range.  Loss: mean squared error.  Weights
are kept in FP32 (Adam) and used as BF16; all products accumulate in FP32 on the
matrix cores.  torch only allocates the buffers and names the stream.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _capi
from .propagator import _ptr, _stream

NOUT = 9


@dataclass(frozen=True)
class NetSpec:
    cin: int = 192
    length: int = 1024
    channels: int = 64
    ks: int = 9
    stride: int = 4
    pad: int = 4
    nblocks: int = 4
    hidden: int = 128
    nout: int = NOUT

    def convs(self):
        """[(cin, cout, lin, lout)] per convolution block."""
        out, c, L = [], self.cin, self.length
        for _ in range(self.nblocks):
            lo = (L + 2 * self.pad - self.ks) // self.stride + 1
            if lo < 1:
                raise ValueError("input too short for the convolution stack")
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

    def flops_per_example(self):
        """Multiply-adds × 2 of one forward pass (the backward is twice this)."""
        f = sum(2 * ci * co * self.ks * lo for ci, co, _, lo in self.convs())
        return f + 2 * self.flat * self.hidden + 2 * self.hidden * self.nout


def he_init(spec: NetSpec, seed=0):
    """He-normal weights (output layer 1/sqrt(fan_in)), zero biases, float32 numpy."""
    rng = np.random.default_rng(seed)
    p = {}
    for k, shp in spec.shapes().items():
        if k[0] in "bc":
            p[k] = np.zeros(shp, np.float32)
        else:
            gain = 1.0 if k == "F2" else 2.0
            p[k] = (rng.standard_normal(shp) * np.sqrt(gain / shp[1])).astype(np.float32)
    return p


def _bf(shape, dev):
    return torch.empty(shape, dtype=torch.bfloat16, device=dev)


def _f(shape, dev):
    return torch.empty(shape, dtype=torch.float32, device=dev)


class InitNet:
    """Network state on one device: FP32 master weights, their BF16 copies, Adam
    moments, gradients, and per-batch-size activation buffers."""

    def __init__(self, spec: NetSpec, params=None, seed=0, device="cuda"):
        self.spec = spec
        self.dev = torch.device(device)
        self.L = _capi.lib()
        p = he_init(spec, seed) if params is None else params
        # every parameter lives in one flat buffer (offsets aligned to 64 values,
        # so each view keeps 16-byte alignment for the vector loads), and so do its
        # BF16 copy, gradient and Adam moments: one Adam launch updates them all
        shapes = spec.shapes()
        offs, n = {}, 0
        for k, shp in shapes.items():
            offs[k] = n
            n += (int(np.prod(shp)) + 63) // 64 * 64
        self.nflat = n
        self.flat_w = torch.zeros(n, device=self.dev)
        self.flat_wb = torch.zeros(n, dtype=torch.bfloat16, device=self.dev)
        self.flat_g = torch.zeros(n, device=self.dev)
        self.flat_m = torch.zeros(n, device=self.dev)
        self.flat_v = torch.zeros(n, device=self.dev)
        view = lambda flat, k: flat[offs[k]:offs[k] + int(np.prod(shapes[k]))].view(shapes[k])
        self.w, self.wb, self.g, self.m, self.v = {}, {}, {}, {}, {}
        for k, shp in shapes.items():
            a = np.asarray(p[k], np.float32)
            if a.shape != tuple(shp):
                raise ValueError(f"parameter {k}: shape {a.shape}, expected {shp}")
            self.w[k] = view(self.flat_w, k)
            self.w[k].copy_(torch.as_tensor(a))
            self.wb[k], self.g[k] = view(self.flat_wb, k), view(self.flat_g, k)
            self.m[k], self.v[k] = view(self.flat_m, k), view(self.flat_v, k)
        self.t = 0
        self._bufs = {}
        self._graphs = {}
        self.sync_bf16()

    # ---- helpers -------------------------------------------------------------
    def _c(self, rc, what):
        _capi.check(rc, what)

    def sync_bf16(self):
        self._c(self.L.efm_net_cast_bf16(_ptr(self.flat_w), _ptr(self.flat_wb), self.nflat, _stream()),
                "efm_net_cast_bf16")

    def params_numpy(self):
        return {k: w.detach().cpu().numpy().copy() for k, w in self.w.items()}

    def _buffers(self, B):
        b = self._bufs.get(B)
        if b is not None:
            return b
        s, dev = self.spec, self.dev
        b = {"x": _bf((s.cin, B, s.length), dev)}
        for i, (ci, co, L, lo) in enumerate(s.convs()):
            b[f"col{i}"] = _bf((ci * s.ks, B * lo), dev)
            b[f"y{i}"] = _f((co, B * lo), dev)
            b[f"yb{i}"] = _bf((co, B * lo), dev)
            b[f"dy{i}"] = _f((co, B * lo), dev)
            b[f"dp{i}"] = _f((co, B * lo), dev)
            b[f"dpb{i}"] = _bf((co, B * lo), dev)
            if i > 0:
                b[f"dcol{i}"] = _f((ci * s.ks, B * lo), dev)
        b["X"] = _bf((s.flat, B), dev)
        b["z"] = _f((s.hidden, B), dev)
        b["zb"] = _bf((s.hidden, B), dev)
        b["out"] = _f((s.nout, B), dev)
        b["t"] = _f((s.nout, B), dev)
        b["dout"] = _f((s.nout, B), dev)
        b["doutb"] = _bf((s.nout, B), dev)
        b["dz"] = _f((s.hidden, B), dev)
        b["dzp"] = _f((s.hidden, B), dev)
        b["dzpb"] = _bf((s.hidden, B), dev)
        b["dX"] = _f((s.flat, B), dev)
        b["loss"] = torch.zeros(1, dtype=torch.float64, device=dev)
        self._bufs[B] = b
        return b

    def _gemm(self, M, N, K, A, sam, sak, Bm, sbk, sbn, C, ldc, bias=None, Cb=None, epi=0):
        self._c(self.L.efm_net_gemm(M, N, K, _ptr(A), sam, sak, _ptr(Bm), sbk, sbn, _ptr(C), ldc,
                                    _ptr(bias) if bias is not None else None,
                                    _ptr(Cb) if Cb is not None else None, epi, _stream()),
                "efm_net_gemm")

    # ---- passes --------------------------------------------------------------
    def forward_buffers(self, B):
        """Forward pass on the input already in the B-sized buffer ``x``."""
        s, b, L = self.spec, self._buffers(B), self.L
        h = b["x"]
        for i, (ci, co, Lin, lo) in enumerate(s.convs()):
            self._c(L.efm_net_im2col(_ptr(h), _ptr(b[f"col{i}"]), ci, B, Lin, s.ks, s.stride, s.pad,
                                     _stream()), "efm_net_im2col")
            K, N = ci * s.ks, B * lo
            self._gemm(co, N, K, self.wb[f"W{i}"], K, 1, b[f"col{i}"], N, 1, b[f"y{i}"], N,
                       bias=self.w[f"b{i}"], Cb=b[f"yb{i}"], epi=1)
            h = b[f"yb{i}"]
        co, lo = s.channels, s.convs()[-1][3]
        self._c(L.efm_net_flatten(_ptr(h), _ptr(b["X"]), co, B, lo, _stream()), "efm_net_flatten")
        self._gemm(s.hidden, B, s.flat, self.wb["F1"], s.flat, 1, b["X"], B, 1, b["z"], B,
                   bias=self.w["c1"], Cb=b["zb"], epi=1)
        self._gemm(s.nout, B, s.hidden, self.wb["F2"], s.hidden, 1, b["zb"], B, 1, b["out"], B,
                   bias=self.w["c2"], epi=2)
        return b["out"]

    def backward_buffers(self, B):
        """MSE loss against the buffer ``t`` and every parameter gradient into self.g.
        Returns the device loss (a 1-element float64 tensor)."""
        s, b, L = self.spec, self._buffers(B), self.L
        b["loss"].zero_()
        n = s.nout * B
        self._c(L.efm_net_mse(_ptr(b["out"]), _ptr(b["t"]), _ptr(b["dout"]), _ptr(b["doutb"]),
                              _ptr(b["loss"]), n, 1.0 / n, _stream()), "efm_net_mse")
        # output layer: dF2 = dout_b · zbᵀ, dc2 = Σ dout, dz = F2ᵀ · dout_b
        self._gemm(s.nout, s.hidden, B, b["doutb"], B, 1, b["zb"], 1, B, self.g["F2"], s.hidden)
        self._c(L.efm_net_bias_grad(_ptr(b["dout"]), _ptr(self.g["c2"]), s.nout, B, _stream()),
                "efm_net_bias_grad")
        self._gemm(s.hidden, B, s.nout, self.wb["F2"], 1, s.hidden, b["doutb"], B, 1, b["dz"], B)
        self._c(L.efm_net_relu_backward(_ptr(b["dz"]), _ptr(b["z"]), _ptr(b["dzp"]), _ptr(b["dzpb"]),
                                        s.hidden * B, _stream()), "efm_net_relu_backward")
        self._gemm(s.hidden, s.flat, B, b["dzpb"], B, 1, b["X"], 1, B, self.g["F1"], s.flat)
        self._c(L.efm_net_bias_grad(_ptr(b["dzp"]), _ptr(self.g["c1"]), s.hidden, B, _stream()),
                "efm_net_bias_grad")
        self._gemm(s.flat, B, s.hidden, self.wb["F1"], 1, s.flat, b["dzpb"], B, 1, b["dX"], B)
        convs = s.convs()
        co, lo = convs[-1][1], convs[-1][3]
        last = len(convs) - 1
        self._c(L.efm_net_unflatten(_ptr(b["dX"]), _ptr(b[f"dy{last}"]), co, B, lo, _stream()),
                "efm_net_unflatten")
        for i in range(last, -1, -1):
            ci, co, Lin, lo = convs[i]
            N, K = B * lo, ci * s.ks
            self._c(L.efm_net_relu_backward(_ptr(b[f"dy{i}"]), _ptr(b[f"y{i}"]), _ptr(b[f"dp{i}"]),
                                            _ptr(b[f"dpb{i}"]), co * N, _stream()),
                    "efm_net_relu_backward")
            # dW = dpre_b · colᵀ, db = Σ dpre
            self._gemm(co, K, N, b[f"dpb{i}"], N, 1, b[f"col{i}"], 1, N, self.g[f"W{i}"], K)
            self._c(L.efm_net_bias_grad(_ptr(b[f"dp{i}"]), _ptr(self.g[f"b{i}"]), co, N, _stream()),
                    "efm_net_bias_grad")
            if i > 0:
                # dcol = Wᵀ · dpre_b, then the transposed im2col into the previous layer
                self._gemm(K, N, co, self.wb[f"W{i}"], 1, K, b[f"dpb{i}"], N, 1, b[f"dcol{i}"], N)
                self._c(L.efm_net_col2im(_ptr(b[f"dcol{i}"]), _ptr(b[f"dy{i - 1}"]), ci, B, Lin, s.ks,
                                         s.stride, s.pad, _stream()), "efm_net_col2im")
        return b["loss"]

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam step over every parameter (the alignment gaps hold zeros and stay zero)."""
        self.t += 1
        c1, c2 = 1.0 / (1 - b1 ** self.t), 1.0 / (1 - b2 ** self.t)
        self._c(self.L.efm_net_adam(_ptr(self.flat_w), _ptr(self.flat_g), _ptr(self.flat_m),
                                    _ptr(self.flat_v), _ptr(self.flat_wb), self.nflat, lr, b1, b2, eps,
                                    c1, c2, _stream()), "efm_net_adam")

    # ---- data ------------------------------------------------------------------
    def load_batch(self, data, targets, idx):
        """Assemble a mini-batch from a (N, cin, length) BF16 dataset and (N, nout)
        targets on the device; idx is a device int32 tensor."""
        s = self.spec
        B = int(idx.numel())
        if data.dtype != torch.bfloat16 or tuple(data.shape[1:]) != (s.cin, s.length):
            raise ValueError(f"dataset must be (N, {s.cin}, {s.length}) bfloat16")
        if targets.shape != (data.shape[0], s.nout) or idx.dtype != torch.int32:
            raise ValueError("targets must be (N, nout) float32 and idx int32")
        b = self._buffers(B)
        self._c(self.L.efm_net_gather_batch(_ptr(data), _ptr(targets), data.shape[0], _ptr(idx), B,
                                            s.cin, s.length, s.nout, _ptr(b["x"]), _ptr(b["t"]),
                                            _stream()), "efm_net_gather_batch")
        return B

    def set_input(self, x):
        """Copy a (cin, B, length) BF16 device tensor into the B-sized input buffer."""
        B = x.shape[1]
        self._buffers(B)["x"].copy_(x)
        return B

    def train_step(self, data, targets, idx, lr=1e-3, graph=False):
        """One Adam step on the mini-batch `idx`.  graph=True replays a hipGraph of
        the batch assembly, forward and backward passes (captured on first use for
        this dataset and batch size; the index list is copied into the graph's
        own buffer), then launches Adam."""
        if not graph:
            B = self.load_batch(data, targets, idx)
            self.forward_buffers(B)
            loss = self.backward_buffers(B)
            self.adam(lr)
            return loss
        B = int(idx.numel())
        key = (B, data.data_ptr(), tuple(data.shape), targets.data_ptr(), tuple(targets.shape))
        gr = self._graphs.get(key)
        if gr is None:
            static_idx = idx.clone()
            self.load_batch(data, targets, static_idx)          # eager warm-up: buffers exist
            self.forward_buffers(B)
            self.backward_buffers(B)
            torch.cuda.synchronize(self.dev)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.load_batch(data, targets, static_idx)
                self.forward_buffers(B)
                loss = self.backward_buffers(B)
            gr = self._graphs[key] = (g, static_idx, loss)
        g, static_idx, loss = gr
        static_idx.copy_(idx)
        g.replay()
        self.adam(lr)
        return loss

    def predict(self, data, idx=None, batch=256):
        """(N, nout) float32 outputs for the examples ``idx`` (default all) of a dataset."""
        N = data.shape[0]
        idx = torch.arange(N, dtype=torch.int32, device=self.dev) if idx is None else idx
        dummy = torch.zeros((N, self.spec.nout), device=self.dev)
        outs = []
        for i in range(0, idx.numel(), batch):
            part = idx[i:i + batch]
            B = self.load_batch(data, dummy, part)
            outs.append(self.forward_buffers(B).t().clone())
        return torch.cat(outs)

    def evaluate(self, data, targets, idx=None, batch=256):
        """Mean squared error over examples ``idx`` (float, synchronises)."""
        pred = self.predict(data, idx, batch)
        t = targets if idx is None else targets[idx.long()]
        return float(((pred.double() - t.double()) ** 2).mean())


def pack_gathers(gathers, length, data, s0):
    """Write normalised BF16 network inputs of gathers (S, 3, ntmax, nrec) into
    rows s0.. of a (N, 3·nrec, length) dataset."""
    S, nc, nt, nr = gathers.shape
    if data.dtype != torch.bfloat16 or tuple(data.shape[1:]) != (nc * nr, length) or length > nt:
        raise ValueError("dataset shape does not match the gathers")
    if s0 < 0 or s0 + S > data.shape[0]:
        raise ValueError("dataset rows out of range")
    _capi.check(_capi.lib().efm_net_pack_gathers(_ptr(gathers), S, nc, nt, nr, length, _ptr(data), s0,
                                                 _stream()), "efm_net_pack_gathers")
