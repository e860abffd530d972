"""Small seeded problems shared by the CPU and GPU tests."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from oracle import elastic_fp64 as O


@dataclass
class Tiny:
    nx: int
    ny: int
    nz: int
    W: int
    free_top: bool
    dx: float
    dt: float
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    lam: np.ndarray
    mu: np.ndarray
    cp: dict
    src: tuple
    m6: np.ndarray
    wavelet: np.ndarray
    rec: np.ndarray
    nt: int

    @property
    def c(self):
        return self.dt / self.dx

    def medium(self):
        return O.Medium(self.lam, self.mu, self.rho)

    def cpml(self):
        return O.Cpml.from_dict(self.cp)


def tiny(n=(14, 13, 12), nt=40, W=3, free_top=False, seed=0, hetero=True, f0=None,
         src=None, nrec=6):
    """Random heterogeneous tiny problem: (nx, ny, nz) = n."""
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    dx = 10.0
    shape = (ny, nz, nx)
    if hetero:
        vp = (3000 + 400 * rng.standard_normal(shape)).astype(np.float32)
        vs = (vp / (1.7 + 0.1 * rng.random(shape))).astype(np.float32)
        rho = (2200 + 200 * rng.random(shape)).astype(np.float32)
    else:
        vp = np.full(shape, 3000, np.float32)
        vs = np.full(shape, 1730, np.float32)
        rho = np.full(shape, 2400, np.float32)
    lam, mu = O.lame_from_velocity(vp, vs, rho)
    vmax = float(vp.max())
    dt = 0.4 * dx / vmax
    f0 = f0 if f0 is not None else 0.12 / dt / 8
    cp = O.cpml_coefficients(W, f0, dt, dx, vmax, nx, ny, nz, free_top=free_top)
    if src is None:
        src = (nx // 2, ny // 2, nz // 2)
    m6 = rng.uniform(-1, 1, 6)
    wavelet = O.ricker(f0, dt, nt)
    rx = rng.integers(2, nx - 2, nrec)
    ry = rng.integers(2, ny - 2, nrec)
    rz = rng.integers(2, nz - 2, nrec)
    rec = np.stack([rx, ry, rz], 1).astype(np.int32)
    return Tiny(nx, ny, nz, W, free_top, dx, dt, vp, vs, rho, lam, mu, cp, src, m6, wavelet, rec, nt)


def run_forward(p: Tiny, m6=None, lam=None, mu=None, rho=None, history=None, threads=1):
    med = O.Medium(p.lam if lam is None else lam, p.mu if mu is None else mu,
                   p.rho if rho is None else rho)
    g, F = O.forward(med, p.cpml(), p.c, p.src, p.m6 if m6 is None else m6, p.wavelet,
                     p.rec, p.nt, threads=threads, history=history)
    return g, F


def rel_l2(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
