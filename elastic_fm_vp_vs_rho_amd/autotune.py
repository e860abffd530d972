"""Stencil tile autotuner.

For each sweep kernel (0 stress, 1 velocity, 2 adjoint velocity, 3 adjoint
stress) every compiled tile variant (bx × bz threads, ny cells per thread along
y, block-to-XCD mapping) is first run once on a live state and its output compared
with the output of the reference tile; a candidate that does not reproduce it
(relative max-norm difference above 1e-6 on any field) is rejected before it is
timed.  Survivors are timed with HIP events on the library's stream; the
fastest is installed with efm_set_tile.  Choices are cached per
(device arch, grid, shot count) in a JSON file.
"""
from __future__ import annotations

import ctypes
import json
import os
import sys

import torch

from . import _capi
from .propagator import Wavefield, _ptr, _stream, padded_components

KERNELS = ("stress", "velocity", "adj_velocity", "adj_stress")
REF_TILE = (64, 4, 1, 0)          # the parity reference every candidate is compared with
DEFAULT_TILE = (128, 1, 1, 16)    # what the library installs until a tile is set
CACHE = os.path.join(os.path.expanduser("~"), ".cache", "efm_tiles.json")


def variants():
    L = _capi.lib()
    n = L.efm_tile_variants(None, 0)
    buf = (ctypes.c_int * (4 * n))()
    L.efm_tile_variants(buf, n)
    return [tuple(buf[4 * i:4 * i + 4]) for i in range(n)]


def shots_per_dispatch(grid):
    """Most shots a batch may hold on `grid`: each sweep runs all shots of a
    batch in one dispatch of fewer than 2^32 work-items, whichever compiled
    tile is installed (the library refuses larger batches)."""
    worst = 0
    for bx, bz, ny, mp in variants():
        nyc = -(-(grid.ny - 4) // ny)
        if mp == 16:    # compact interior plane
            per = -(-(grid.nx - 4) * (grid.nz - 4) // (bx * bz)) * bx * bz
        else:
            per = -(-grid.nx // bx) * bx * -(-grid.nz // bz) * bz
        worst = max(worst, per * nyc)
    return max(1, ((1 << 32) - 1 - 8 * 512) // worst)


def get_tile(k):
    v = [ctypes.c_int() for _ in range(4)]
    _capi.check(_capi.lib().efm_get_tile(k, *[ctypes.byref(x) for x in v]), "efm_get_tile")
    return tuple(x.value for x in v)


def set_tile(k, tile):
    _capi.check(_capi.lib().efm_set_tile(k, *tile), "efm_set_tile")


def _key(prop):
    props = torch.cuda.get_device_properties(prop.lam.device)
    arch = getattr(props, "gcnArchName", props.name)
    g = prop.grid
    return f"{arch}:{g.nx}x{g.ny}x{g.nz}:S{prop.shots.S}:w{','.join(map(str, g.widths))}"


def _relmax(a, b):
    d = float((a.double() - b.double()).abs().max())
    s = float(b.double().abs().max())
    return d / s if s > 0 else d


class _Bench:
    """Holds a live forward state and an adjoint set-up for timing kernels."""

    def __init__(self, prop, warm_steps=60, n0=None):
        self.prop = prop
        self.L = _capi.lib()
        s = prop.shots
        nt = s.ntmax
        self.n0 = min(warm_steps, nt - 3) if n0 is None else n0
        self.d, self.m, self.sh = prop.dims(), prop.medium(), s.struct()
        _, st = prop.forward(nsteps=self.n0)
        self.base = st
        self.work = st.clone()
        dev = prop.lam.device
        g = prop.grid
        self.states = [st.clone(), Wavefield(g, s.S, dev), Wavefield(g, s.S, dev)]
        self.arr = (ctypes.POINTER(_capi.EfmState) * 3)(*[ctypes.pointer(w.struct()) for w in self.states])
        _capi.check(self.L.efm_forward_states(ctypes.byref(self.d), ctypes.byref(self.m),
                                              ctypes.byref(self.sh), self.arr, None, self.n0, 2,
                                              _stream()), "efm_forward_states")
        self.mbar = padded_components(3, g.N, dev)
        self.bbar = padded_components(3, g.N, dev)
        _capi.check(self.L.efm_derived_medium(ctypes.byref(self.d), ctypes.byref(self.m),
                                              _ptr(self.mbar), _ptr(self.bbar), _stream()))
        gen = torch.Generator(device=dev).manual_seed(1)
        self.resid = torch.randn((s.S, 3, nt, s.nrec), device=dev, generator=gen) * 1e-3
        self.adj0 = Wavefield(g, s.S, dev)
        for k in range(9):
            self.adj0.fields[k].copy_(torch.randn(self.adj0.fields[k].shape, device=dev, generator=gen))
        self.adj = self.adj0.clone()
        self.adj2 = Wavefield(g, s.S, dev, psi_only_of=self.adj)
        self.acc = padded_components(s.S * 8, g.N, dev)
        self.gm6 = torch.zeros((s.S, 6), dtype=torch.float64, device=dev)

    def run(self, k):
        """Run kernel k once from a fixed input; return output tensors."""
        L, d, m, sh = self.L, self.d, self.m, self.sh
        if k in (0, 1):
            self.work.copy_(self.base)
            w = ctypes.byref(self.work.struct())
            fn = L.efm_stress_step if k == 0 else L.efm_velocity_step
            _capi.check(fn(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh), w, w, self.n0, _stream()))
            return [self.work.fields, self.work.psi_x, self.work.psi_y, self.work.psi_z]
        self.adj.copy_(self.adj0)
        self.acc.zero_()
        self.gm6.zero_()
        self.adjoint_once()
        return [self.adj.fields, self.acc, self.adj.psi_x, self.adj.psi_y, self.adj.psi_z]

    def adjoint_once(self):
        _capi.check(self.L.efm_adjoint(ctypes.byref(self.d), ctypes.byref(self.m), _ptr(self.mbar),
                                       _ptr(self.bbar), ctypes.byref(self.sh), self.arr,
                                       ctypes.byref(self.adj.struct()), ctypes.byref(self.adj2.struct()),
                                       _ptr(self.resid), _ptr(self.acc), _ptr(self.gm6), self.n0,
                                       self.n0 + 2, 0, _stream()), "efm_adjoint")

    def time(self, k, reps):
        L, d, m, sh = self.L, self.d, self.m, self.sh
        if k in (0, 1):
            w = ctypes.byref(self.work.struct())
            fn = L.efm_stress_step if k == 0 else L.efm_velocity_step
            call = lambda: fn(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh), w, w, self.n0, _stream())
        else:
            call = self.adjoint_once
        call()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            call()
        b.record()
        torch.cuda.synchronize()
        per = a.elapsed_time(b) / reps
        return per / 2 if k in (2, 3) else per      # an adjoint segment call runs 2 steps


def tune(prop, kernels=(0, 1, 2, 3), reps=20, tol=1e-6, use_cache=True, verbose=False):
    """Choose and install the fastest parity-clean tile per kernel.  Returns
    {kernel name: {"tile": tile, "ms": {tile: ms or "rejected"}}}."""
    key = _key(prop)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        # a cache written by another build may name tiles this library lacks: re-tune then
        compiled = set(variants())
        entry = cache[key]
        if all(KERNELS[k] in entry and tuple(entry[KERNELS[k]]["tile"]) in compiled for k in kernels):
            for k in kernels:
                set_tile(k, tuple(entry[KERNELS[k]]["tile"]))
            return entry
    B = _Bench(prop)
    report = {}
    vs = variants()
    for k in kernels:
        others = {j: get_tile(j) for j in range(4)}
        set_tile(k, REF_TILE)
        ref = [t.clone() for t in B.run(k)]
        times = {}
        for tile in vs:
            set_tile(k, tile)
            out = B.run(k)
            err = max(_relmax(o, r) for o, r in zip(out, ref))
            if err > tol:
                times[str(tile)] = f"rejected (rel diff {err:.1e})"
                continue
            times[str(tile)] = B.time(k, reps)
        ok = {t: v for t, v in times.items() if not isinstance(v, str)}
        best = min(ok, key=ok.get)
        tile = tuple(int(x) for x in best.strip("()").split(","))
        set_tile(k, tile)
        for j, t in others.items():
            if j != k:
                set_tile(j, t)
        report[KERNELS[k]] = {"tile": tile, "ms": times}
        if verbose:
            print(f"[autotune] {KERNELS[k]}: best {tile} {ok[best]:.4f} ms; "
                  + ", ".join(f"{t}: {v if isinstance(v, str) else f'{v:.4f}'}" for t, v in times.items()),
                  file=sys.stderr)
    if use_cache:
        cache[key] = report
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            with open(CACHE, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    return report
