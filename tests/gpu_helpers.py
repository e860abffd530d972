"""Helpers for GPU tests: build device problems from oracle-side tiny problems
and convert between the oracle's full-volume ψ and the device slab layout."""
from __future__ import annotations

import numpy as np
import torch

from oracle import elastic_fp64 as O

PSI_X = ("vx_x", "vy_x", "vz_x", "txx_x", "txy_x", "txz_x")
PSI_Y = ("vx_y", "vy_y", "vz_y", "txy_y", "tyy_y", "tyz_y")
PSI_Z = ("vx_z", "vy_z", "vz_z", "txz_z", "tyz_z", "tzz_z")


def grid_of(p):
    from elastic_fm_vp_vs_rho_amd.propagator import Grid
    W = p.W
    return Grid(p.nx, p.ny, p.nz, p.dx, p.dt, (W, W, W, W, 0 if p.free_top else W, W))


def shots_of(p, grid, m6=None, nt=None, wavelet=None):
    from elastic_fm_vp_vs_rho_amd.propagator import ShotBatch
    nt = p.nt if nt is None else nt
    w = p.wavelet if wavelet is None else wavelet
    return ShotBatch(grid, [p.src], [p.m6 if m6 is None else m6], [w], [nt], [p.rec], p.cp)


def propagator_of(p, lam=None, mu=None, rho=None, m6=None):
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator
    grid = grid_of(p)
    dev = torch.device("cuda")
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
    return Propagator(grid, f(p.lam if lam is None else lam), f(p.mu if mu is None else mu),
                      f(p.rho if rho is None else rho), shots_of(p, grid, m6))


def slab_axis(n, lo, hi):
    return np.concatenate([np.arange(lo), np.arange(n - hi, n)]).astype(int)


def state_to_device(F, wf):
    """Oracle Fields (full-volume ψ) -> device Wavefield (shot 0)."""
    g = wf.grid
    w = g.widths
    for i, k in enumerate(O.FIELDS):
        wf.fields[i, 0].copy_(torch.as_tensor(F.f[k].astype(np.float32)))
    xs, ys, zs = slab_axis(g.nx, w[0], w[1]), slab_axis(g.ny, w[2], w[3]), slab_axis(g.nz, w[4], w[5])
    for c, k in enumerate(PSI_X):
        wf.psi_x[0, c].copy_(torch.as_tensor(F.psi[k][:, :, xs].astype(np.float32)))
    for c, k in enumerate(PSI_Y):
        wf.psi_y[0, c].copy_(torch.as_tensor(F.psi[k][ys].astype(np.float32)))
    for c, k in enumerate(PSI_Z):
        wf.psi_z[0, c].copy_(torch.as_tensor(F.psi[k][:, zs].astype(np.float32)))


def state_from_device(wf, shape):
    F = O.Fields(shape)
    g = wf.grid
    w = g.widths
    for i, k in enumerate(O.FIELDS):
        F.f[k][...] = wf.fields[i, 0].double().cpu().numpy()
    xs, ys, zs = slab_axis(g.nx, w[0], w[1]), slab_axis(g.ny, w[2], w[3]), slab_axis(g.nz, w[4], w[5])
    for c, k in enumerate(PSI_X):
        F.psi[k][:, :, xs] = wf.psi_x[0, c].double().cpu().numpy()
    for c, k in enumerate(PSI_Y):
        F.psi[k][ys] = wf.psi_y[0, c].double().cpu().numpy()
    for c, k in enumerate(PSI_Z):
        F.psi[k][:, zs] = wf.psi_z[0, c].double().cpu().numpy()
    return F


def random_state(p, rng):
    """Random interior v, τ and random ψ inside the CPML slabs (zero elsewhere,
    as in any real run), float32-representable."""
    shape = (p.ny, p.nz, p.nx)
    F = O.Fields(shape)
    I = (slice(2, p.ny - 2), slice(2, p.nz - 2), slice(2, p.nx - 2))
    for k in O.FIELDS:
        a = np.zeros(shape)
        a[I] = rng.standard_normal(a[I].shape)
        F.f[k][...] = a.astype(np.float32)
    W = p.W
    wz = 0 if p.free_top else W
    X = np.zeros(p.nx, bool); X[:W] = True; X[p.nx - W:] = True
    Y = np.zeros(p.ny, bool); Y[:W] = True; Y[p.ny - W:] = True
    Z = np.zeros(p.nz, bool); Z[:wz] = True; Z[p.nz - W:] = True
    masks = {"x": X[None, None, :], "y": Y[:, None, None], "z": Z[None, :, None]}
    for k in O.PSI_V + O.PSI_T:
        a = 0.1 * rng.standard_normal(shape) * masks[k[-1]]
        F.psi[k][...] = a.astype(np.float32)
    return F
