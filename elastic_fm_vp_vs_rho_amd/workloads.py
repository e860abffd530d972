"""Seeded synthetic workloads (models, geometry, wavelets) for the benchmark
configurations.  Host numpy only; no device work here.

C1  plain forward: 64³, dx = 10 m, homogeneous Vp 3000 / Vs 1730 / ρ 2400,
    dt at Courant 0.5, 10-cell CPML on all faces, MT source at the centre
    (6 components U[-1, 1], seed 1, normalised by the tensor's Frobenius norm),
    10 Hz Ricker, 1000 steps, 16 receivers on a 4×4 grid one level below the
    top CPML.
C2  forward + adjoint gradient: 128³, dx = 10 m, three horizontal layers
    (Vp 2500/3200/4000, Vs = Vp/1.73, ρ = 310·Vp^0.25 Gardner), a +10 % Vp and Vs
    Gaussian anomaly (σ = 8 cells) at the centre in the "true" model, gradient
    on the layered model; one MT source (seed 2) at depth z = 100, 10 Hz Ricker,
    2000 steps, 64 receivers (8×8).
C3  batched: 256×256×128 random smooth layered model (seed 4), 32 sources at
    z ∈ U[20, 100], f0 ∈ U[5, 15] Hz, nt ∈ U[1500, 3000], 128 receivers.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import model as M


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dt: float
    W: int
    free_top: bool
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    src: np.ndarray          # (S, 3) x, y, z
    m6: np.ndarray           # (S, 6)
    f0: np.ndarray           # (S,)
    nt: np.ndarray           # (S,)
    rec: np.ndarray          # (S, nrec, 3)
    vmax: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def S(self):
        return self.src.shape[0]

    @property
    def ntmax(self):
        return int(self.nt.max())

    def wavelets(self):
        w = np.zeros((self.S, self.ntmax), np.float32)
        for s in range(self.S):
            w[s, : self.nt[s]] = M.ricker(float(self.f0[s]), self.dt, int(self.nt[s]))
        return w

    def cpml(self):
        """Per-shot CPML profiles (they depend on each shot's f0)."""
        out = {k: [] for k in ("ax", "bx", "ay", "by", "az", "bz")}
        for s in range(self.S):
            c = M.cpml_coefficients(self.W, float(self.f0[s]), self.dt, self.dx, self.vmax,
                                    self.nx, self.ny, self.nz, free_top=self.free_top)
            for k in out:
                out[k].append(c[k])
        return {k: np.stack(v) for k, v in out.items()}

    def lame(self):
        return M.lame_from_velocity(self.vp, self.vs, self.rho)

    def cell_steps(self):
        return float(self.nx * self.ny * self.nz) * float(self.nt.sum())


def normalized_mt(rng):
    m = rng.uniform(-1.0, 1.0, 6)
    fro = np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2 + 2 * (m[3] ** 2 + m[4] ** 2 + m[5] ** 2))
    return m / fro


def surface_grid(nx, ny, W, z, nline_x, nline_y):
    """Receivers x-inner, y-outer like the reference (M:2851-2861)."""
    xx = np.linspace(W, nx - W - 1, nline_x).astype(int)
    yy = np.linspace(W, ny - W - 1, nline_y).astype(int)
    return np.array([(x, y, z) for y in yy for x in xx], np.int64)


def courant_dt(dx, vmax, courant=0.5):
    return courant * dx / vmax


def c1(n=64, nt=1000):
    dx, W = 10.0, 10
    shape = (n, n, n)
    vp = np.full(shape, 3000.0, np.float32)
    vs = np.full(shape, 1730.0, np.float32)
    rho = np.full(shape, 2400.0, np.float32)
    vmax = 3000.0
    dt = courant_dt(dx, vmax)
    rng = np.random.default_rng(1)
    m6 = normalized_mt(rng)
    src = np.array([[n // 2, n // 2, n // 2]])
    rec = surface_grid(n, n, W, W, 4, 4)[None]
    return Workload("C1 plain forward 64^3 x 1000 steps", n, n, n, dx, dt, W, False, vp, vs, rho, src,
                    m6[None], np.array([10.0]), np.array([nt]), rec, vmax)


def layered(n, anomaly):
    nx = ny = nz = n
    z = np.arange(nz)
    vp_z = np.where(z < nz // 3, 2500.0, np.where(z < 2 * nz // 3, 3200.0, 4000.0))
    vp = np.broadcast_to(vp_z[None, :, None], (ny, nz, nx)).astype(np.float64).copy()
    vs = vp / 1.73
    if anomaly:
        c = (n - 1) / 2.0
        yy, zz, xx = np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij")
        gss = np.exp(-((xx - c) ** 2 + (yy - c) ** 2 + (zz - c) ** 2) / (2 * 8.0 ** 2))
        vp = vp * (1 + 0.10 * gss)
        vs = vs * (1 + 0.10 * gss)
    rho = 310.0 * vp ** 0.25
    return vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32)


def c2(n=128, nt=2000, anomaly=False):
    dx, W = 10.0, 10
    vp, vs, rho = layered(n, anomaly)
    vmax = 4000.0 * 1.1
    dt = courant_dt(dx, vmax)
    rng = np.random.default_rng(2)
    m6 = normalized_mt(rng)
    sx, sy = n // 2 + int(rng.integers(-8, 9)), n // 2 + int(rng.integers(-8, 9))
    src = np.array([[sx, sy, min(100, n - W - 4)]])
    rec = surface_grid(n, n, W, W, 8, 8)[None]
    name = f"C2 fwd+adjoint gradient {n}^3 x {nt} steps" + (" (true model)" if anomaly else "")
    return Workload(name, n, n, n, dx, dt, W, False, vp, vs, rho, src, m6[None], np.array([10.0]),
                    np.array([nt]), rec, vmax)


def c3(nx=256, ny=256, nz=128, S=32, nt_range=(1500, 3000), nrec=128, seed=4):
    dx, W = 10.0, 10
    rng = np.random.default_rng(seed)
    nlay = 6
    bounds = np.sort(rng.integers(10, nz - 10, nlay - 1))
    vps = np.sort(rng.uniform(2000.0, 4500.0, nlay))
    vp_z = np.zeros(nz)
    lo = 0
    for i, b in enumerate(list(bounds) + [nz]):
        vp_z[lo:b] = vps[i]
        lo = b
    k = np.exp(-0.5 * (np.arange(-8, 9) / 3.0) ** 2)
    vp_z = np.convolve(np.pad(vp_z, 8, mode="edge"), k / k.sum(), mode="valid")
    vp = np.broadcast_to(vp_z[None, :, None], (ny, nz, nx)).astype(np.float32).copy()
    vs = (vp / 1.73).astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    vmax = float(vp.max())
    dt = courant_dt(dx, vmax)
    src = np.stack([rng.integers(W + 4, nx - W - 4, S), rng.integers(W + 4, ny - W - 4, S),
                    rng.integers(20, 101, S)], 1)
    m6 = np.stack([normalized_mt(rng) for _ in range(S)])
    f0 = rng.uniform(5.0, 15.0, S)
    nt = rng.integers(nt_range[0], nt_range[1] + 1, S)
    side = int(np.sqrt(nrec))
    rec1 = surface_grid(nx, ny, W, W, nrec // side, side)
    rec = np.broadcast_to(rec1[None], (S,) + rec1.shape).copy()
    return Workload(f"C3 batched {nx}x{ny}x{nz}, {S} shots", nx, ny, nz, dx, dt, W, False, vp, vs,
                    rho, src, m6, f0, nt, rec, vmax)


def c5(n=96, S=20000, nt=1024, nrec=64, f0=10.0, seed=5):
    """C5 initializer data: one 96³ random layered model (seed 5), S sources
    with uniform random location (x, y, z in [W+4, n-W-4)) and normalised MT,
    10 Hz Ricker, nt steps, 8×8 receivers one level below the top CPML."""
    dx, W = 10.0, 10
    rng = np.random.default_rng(seed)
    nlay = 5
    bounds = np.sort(rng.choice(np.arange(12, n - 12), nlay - 1, replace=False))
    vps = np.sort(rng.uniform(2500.0, 4500.0, nlay))
    vp_z = np.zeros(n)
    lo = 0
    for i, b in enumerate(list(bounds) + [n]):
        vp_z[lo:b] = vps[i]
        lo = b
    vp = np.broadcast_to(vp_z[None, :, None], (n, n, n)).astype(np.float32).copy()
    vs = (vp / 1.75).astype(np.float32)
    rho = M.rho_nafe_drake(vp).astype(np.float32)
    vmax = float(vp.max())
    dt = courant_dt(dx, vmax)
    lo_, hi_ = W + 4, n - W - 4
    src = rng.integers(lo_, hi_, (S, 3))
    m6 = np.stack([normalized_mt(rng) for _ in range(S)])
    side = int(np.sqrt(nrec))
    rec1 = surface_grid(n, n, W, W, nrec // side, side)
    rec = np.broadcast_to(rec1[None], (S,) + rec1.shape)
    return Workload(f"C5 initializer data {n}^3, {S} shots x {nt} steps", n, n, n, dx, dt, W, False,
                    vp, vs, rho, src, m6, np.full(S, f0), np.full(S, nt), rec, vmax,
                    extra={"src_range": (lo_, hi_)})


def c4(nx=200, ny=200, nz=100, S=16, nt=2500, nrec=64, f0=10.0, anomalies=True, seed=6):
    """C4 FWI + MT: smoothly layered model (Vp 2600 -> 4200 m/s with depth,
    Vs = Vp/1.73, Gardner ρ = 310·Vp^0.25) with two Gaussian anomalies (+8 % and
    -8 % on Vp and Vs, so ±2 % on ρ through Gardner; σ = 8 cells) in the true
    model; S sources at 0.3-0.7 of the depth
    with normalised MT (seed 6), 10 Hz Ricker, nt steps, 8×8 receivers one
    level below the top CPML; 10-cell CPML on all faces."""
    dx, W = 10.0, 10
    rng = np.random.default_rng(seed)
    vp_z = np.interp(np.arange(nz), [0, nz // 4, nz // 2, 3 * nz // 4, nz - 1],
                     [2600.0, 2900.0, 3300.0, 3800.0, 4200.0])
    vp = np.broadcast_to(vp_z[None, :, None], (ny, nz, nx)).astype(np.float64).copy()
    if anomalies:
        yy, zz, xx = np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij")
        for (cx, cy, cz, amp) in ((0.35 * nx, 0.4 * ny, 0.55 * nz, 0.08), (0.65 * nx, 0.6 * ny, 0.45 * nz, -0.08)):
            vp *= 1 + amp * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) / (2 * 8.0 ** 2))
    vs = vp / 1.73
    rho = 310.0 * vp ** 0.25
    vmax = 4200.0 * 1.1
    dt = courant_dt(dx, vmax)
    mx, my = max(2, min(10, (nx - 2 * W) // 4)), max(2, min(10, (ny - 2 * W) // 4))
    zlo, zhi = max(W + 2, int(0.3 * nz)), min(nz - W - 2, int(0.7 * nz) + 1)
    src = np.stack([rng.integers(W + mx, nx - W - mx, S), rng.integers(W + my, ny - W - my, S),
                    rng.integers(zlo, zhi, S)], 1)
    m6 = np.stack([normalized_mt(rng) for _ in range(S)])
    side = int(np.sqrt(nrec))
    rec1 = surface_grid(nx, ny, W, W, nrec // side, side)
    rec = np.broadcast_to(rec1[None], (S,) + rec1.shape).copy()
    name = f"C4 FWI+MT {nx}x{ny}x{nz}, {S} shots x {nt} steps" + (" (true model)" if anomalies else "")
    return Workload(name, nx, ny, nz, dx, dt, W, False, vp.astype(np.float32), vs.astype(np.float32),
                    rho.astype(np.float32), src, m6, np.full(S, f0), np.full(S, nt), rec, vmax)
