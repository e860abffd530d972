"""FP64 NumPy restatement of the 3-D elastic velocity-stress FDTD path.

TEST INFRASTRUCTURE ONLY (the checker).  Nothing in the product package imports
this module; tests, smoke() and bench.py's cpu_baseline leg do.

Every function cites the reference lines it restates.  Citations are relative to
``Elastic3DGPU_MT_VP_VS_RHO_MPI/mainPropElastic3D_MT.py`` (written ``M:n``).

Conventions (reference ``M:72, 758``): volumes are ``(ny, nz, nx)`` C-order,
x fastest; only the interior ``2 <= x,y,z < n-2`` is updated (``M:54, 77, 152``).
All arithmetic here is float64.  Inputs that the GPU sees in float32 (Lame
parameters, density, CPML profiles, wavelet) are given to the oracle as the same
float32 values, upcast -- so the comparison measures the propagator, not the
input rounding.

Parity pins for each part are listed in ``tests/test_oracle.py``; nothing here is
"parity unpinned" except the scipy-based smoothing, which calls scipy directly.
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

REF_DT = 0.08551979923718353   # M:1254 (hard-coded sample interval)
H_STENCIL = 2                  # M:54

# ψ component names.  Stress sweep (derivatives of v) and velocity sweep
# (derivatives of τ); the suffix is the derivative axis.  Reference names:
# psidVx_dx ... (M:62-65) and psidTauxx_dx ... (M:138-140).
PSI_V = ("vx_x", "vy_x", "vz_x", "vx_y", "vy_y", "vz_y", "vx_z", "vy_z", "vz_z")
PSI_T = ("txx_x", "txy_x", "txz_x", "txy_y", "tyy_y", "tyz_y", "txz_z", "tyz_z", "tzz_z")
FIELDS = ("vx", "vy", "vz", "txx", "tyy", "tzz", "txy", "txz", "tyz")


# ----------------------------------------------------------------------------
# host-side maps (float32, exactly the reference's numpy arithmetic)
# ----------------------------------------------------------------------------
def lame_from_velocity(vp, vs, rho):
    """Model.compute_Lambda_Mu, M:1196-1198 (float32 numpy, same op order)."""
    vp = np.asarray(vp, np.float32); vs = np.asarray(vs, np.float32)
    rho = np.asarray(rho, np.float32)
    mu = vs * vs * rho
    lam = vp * vp * rho - np.float32(2.0) * mu
    return lam, mu


def velocity_from_lame(lam, mu, rho):
    """Model.compute_Vp_Vs, M:1200-1202."""
    vs = np.sqrt(mu / rho)
    vp = np.sqrt((lam + np.float32(2.0) * mu) / rho)
    return vp, vs


def rho_nafe_drake(vp):
    """Model.compute_Rho (Nafe-Drake polynomial, coefficient 0.4712), M:1204-1206."""
    v = vp / 1e3
    return (1.6612 * v - 0.4712 * (v ** 2) + 0.0671 * (v ** 3) - 0.0043 * (v ** 4)
            + 0.000106 * (v ** 5)) * 1e3


def cpml_profile(frequency, dt, di, W, vmax, N):
    """CPML.CPML_i, M:1340-1372: float32 (a, b) profiles of length N."""
    a = np.zeros([N], np.float32)
    b = np.zeros([N], np.float32)
    if W <= 0:                     # no absorbing layer: a = b = 0 everywhere
        return a, b
    R = 1e-6
    Li = W * di
    d0 = -3 * np.log(R) / (2.0 * Li)
    for ix in range(N):
        if ix < W:
            Fi = (W - 1 - ix) * di
            d_i = d0 * vmax * (Fi / Li) * (Fi / Li)
            alpha = np.pi * frequency * ((Li - Fi) / Li)
            b[ix] = np.exp(-(d_i + alpha) * dt)
            a[ix] = (d_i / (d_i + alpha)) * (b[ix] - 1.0)
        if W <= ix < N - W:
            b[ix] = 0.0
            a[ix] = 0.0
        if N - W <= ix < N:
            Fi = (ix - (N - W)) * di
            d_i = d0 * vmax * (Fi / Li) * (Fi / Li)
            alpha = np.pi * frequency * (Li - Fi) / Li
            b[ix] = np.exp(-(d_i + alpha) * dt)
            a[ix] = (d_i / (d_i + alpha)) * (b[ix] - 1.0)
    return a, b


def cpml_coefficients(W, frequency, dt, dx, vmax, nx, ny, nz, free_top=True, dy=None, dz=None):
    """CPML.__init__, M:1298-1307.  ``free_top`` zeroes the z-min profile (M:1305-1306)."""
    dy = dx if dy is None else dy
    dz = dx if dz is None else dz
    ax, bx = cpml_profile(frequency, dt, dx, W, vmax, nx)
    ay, by = cpml_profile(frequency, dt, dy, W, vmax, ny)
    az, bz = cpml_profile(frequency, dt, dz, W, vmax, nz)
    if free_top:
        az[:W] = 0.0
        bz[:W] = 0.0
    return dict(ax=ax, bx=bx, ay=ay, by=by, az=az, bz=bz)


def ricker(fq, dt, nt, fq_exp=None):
    """Wavelet.ricker, M:1264-1272.  The exponential uses the module-global ``fq``
    in the reference (M:1270); ``fq_exp`` reproduces that (default: ``fq``)."""
    fq_exp = fq if fq_exp is None else fq_exp
    t0 = (np.sqrt(2) / fq) / dt
    w = np.zeros(nt, np.float32)
    for t in range(nt):
        w[t] = (1 - 2 * (math.pi * fq * ((t - t0) * dt)) ** 2) * math.exp(
            -(math.pi * fq_exp * ((t - t0) * dt)) ** 2)
    return w


def gaussian(fq, dt, nt):
    """Wavelet.gaussian, M:1274-1281."""
    t0 = (1.0 / fq) / dt
    w = np.zeros(nt, np.float32)
    a = 2 * (np.pi) ** 2 * fq ** 2
    for t in range(nt):
        w[t] = math.exp(-a * ((t - t0) * dt) ** 2)
    return w


# ----------------------------------------------------------------------------
# medium, state
# ----------------------------------------------------------------------------
def _I(n):
    return slice(H_STENCIL, n - H_STENCIL)


def _S(n, d):
    return slice(H_STENCIL + d, n - H_STENCIL + d)


class Medium:
    """Float64 copies of λ, μ, ρ plus the averaged moduli/buoyancies the kernels
    form on the fly: μ̄ harmonic means (M:80-89) and b̄ = 2bb'/(b+b') (M:153-159).
    Derived arrays are defined on the interior and zero elsewhere."""

    def __init__(self, lam, mu, rho):
        self.lam = np.asarray(lam, np.float64)
        self.mu = np.asarray(mu, np.float64)
        self.rho = np.asarray(rho, np.float64)
        ny, nz, nx = self.lam.shape
        self.shape = (ny, nz, nx)
        Y, Z, X = _I(ny), _I(nz), _I(nx)
        mu = self.mu

        def m(dy, dz, dx):
            return mu[_S(ny, dy), _S(nz, dz), _S(nx, dx)]

        with np.errstate(divide="ignore"):
            q = lambda a: 0.25 / a
            self.mu_xz = np.zeros_like(mu); self.mu_xy = np.zeros_like(mu); self.mu_yz = np.zeros_like(mu)
            self.mu_xz[Y, Z, X] = 1.0 / (q(m(0, 0, 0)) + q(m(0, 0, 1)) + q(m(0, 1, 0)) + q(m(0, 1, 1)))
            self.mu_xy[Y, Z, X] = 1.0 / (q(m(0, 0, 0)) + q(m(0, 0, 1)) + q(m(1, 0, 0)) + q(m(1, 0, 1)))
            self.mu_yz[Y, Z, X] = 1.0 / (q(m(0, 0, 0)) + q(m(1, 0, 0)) + q(m(0, 1, 0)) + q(m(1, 1, 0)))
        rho = self.rho

        def r(dy, dz, dx):
            return rho[_S(ny, dy), _S(nz, dz), _S(nx, dx)]

        b = 1.0 / r(0, 0, 0)
        bi, bk, bj = 1.0 / r(0, 0, 1), 1.0 / r(0, 1, 0), 1.0 / r(1, 0, 0)
        self.bbx = np.zeros_like(rho); self.bby = np.zeros_like(rho); self.bbz = np.zeros_like(rho)
        self.bbx[Y, Z, X] = 2.0 * b * bi / (b + bi)
        self.bby[Y, Z, X] = 2.0 * b * bj / (b + bj)
        self.bbz[Y, Z, X] = 2.0 * b * bk / (b + bk)


@dataclass
class Cpml:
    ax: np.ndarray
    bx: np.ndarray
    ay: np.ndarray
    by: np.ndarray
    az: np.ndarray
    bz: np.ndarray

    @classmethod
    def from_dict(cls, d):
        return cls(*(np.asarray(d[k], np.float64) for k in ("ax", "bx", "ay", "by", "az", "bz")))

    @classmethod
    def zeros(cls, nx, ny, nz):
        return cls(np.zeros(nx), np.zeros(nx), np.zeros(ny), np.zeros(ny), np.zeros(nz), np.zeros(nz))


class Fields:
    """The 30-volume state of reference ``Fields`` (M:754-862) without host mirrors
    and borders: v, τ and 18 full-volume ψ (float64)."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.f = {k: np.zeros(shape) for k in FIELDS}
        self.psi = {k: np.zeros(shape) for k in PSI_V + PSI_T}

    def copy(self):
        o = Fields.__new__(Fields)
        o.shape = self.shape
        o.f = {k: v.copy() for k, v in self.f.items()}
        o.psi = {k: v.copy() for k, v in self.psi.items()}
        return o

    def __getattr__(self, k):
        f = self.__dict__.get("f")
        if f is not None and k in f:
            return f[k]
        raise AttributeError(k)


def _cp_views(cp, Y, Z, X):
    return (cp.ax[X][None, None, :], cp.bx[X][None, None, :],
            cp.ay[Y][:, None, None], cp.by[Y][:, None, None],
            cp.az[Z][None, :, None], cp.bz[Z][None, :, None])


def _psi_update(arr, Y, Z, X, a, b, d):
    """ψ = b·ψ + a·Δ in place on the interior window; returns the new window."""
    p = arr[Y, Z, X]
    p *= b
    p += a * d
    return p


# ----------------------------------------------------------------------------
# forward sweeps (M:59-128, M:133-188)
# ----------------------------------------------------------------------------
def stress_sweep(F: Fields, med: Medium, cp: Cpml, c: float, y0=None, y1=None):
    """kernel_Tau, M:59-128, over interior y in [y0, y1)."""
    ny, nz, nx = F.shape
    y0 = H_STENCIL if y0 is None else y0
    y1 = ny - H_STENCIL if y1 is None else y1
    if y1 <= y0:
        return
    Y, Ym, Yp = slice(y0, y1), slice(y0 - 1, y1 - 1), slice(y0 + 1, y1 + 1)
    Z, Zm, Zp = _I(nz), _S(nz, -1), _S(nz, 1)
    X, Xm, Xp = _I(nx), _S(nx, -1), _S(nx, 1)
    f, P = F.f, F.psi
    vx, vy, vz = f["vx"], f["vy"], f["vz"]
    vxc, vyc, vzc = vx[Y, Z, X], vy[Y, Z, X], vz[Y, Z, X]
    ax, bx, ay, by, az, bz = _cp_views(cp, Y, Z, X)
    # raw differences, M:91-100
    dvx_dx = vxc - vx[Y, Z, Xm]
    dvy_dy = vyc - vy[Ym, Z, X]
    dvz_dz = vzc - vz[Y, Zm, X]
    dvx_dz = vx[Y, Zp, X] - vxc
    dvx_dy = vx[Yp, Z, X] - vxc
    dvy_dx = vy[Y, Z, Xp] - vyc
    dvy_dz = vy[Y, Zp, X] - vyc
    dvz_dx = vz[Y, Z, Xp] - vzc
    dvz_dy = vz[Yp, Z, X] - vzc
    # ψ updated first, then used (M:102-112, Q4)
    exx = dvx_dx + _psi_update(P["vx_x"], Y, Z, X, ax, bx, dvx_dx)
    eyx = dvy_dx + _psi_update(P["vy_x"], Y, Z, X, ax, bx, dvy_dx)
    ezx = dvz_dx + _psi_update(P["vz_x"], Y, Z, X, ax, bx, dvz_dx)
    eyy = dvy_dy + _psi_update(P["vy_y"], Y, Z, X, ay, by, dvy_dy)
    exy = dvx_dy + _psi_update(P["vx_y"], Y, Z, X, ay, by, dvx_dy)
    ezy = dvz_dy + _psi_update(P["vz_y"], Y, Z, X, ay, by, dvz_dy)
    ezz = dvz_dz + _psi_update(P["vz_z"], Y, Z, X, az, bz, dvz_dz)
    exz = dvx_dz + _psi_update(P["vx_z"], Y, Z, X, az, bz, dvx_dz)
    eyz = dvy_dz + _psi_update(P["vy_z"], Y, Z, X, az, bz, dvy_dz)
    lam = med.lam[Y, Z, X]
    l2m = lam + 2.0 * med.mu[Y, Z, X]
    # M:114-125
    f["txx"][Y, Z, X] += c * (l2m * exx + lam * (eyy + ezz))
    f["tyy"][Y, Z, X] += c * (l2m * eyy + lam * (exx + ezz))
    f["tzz"][Y, Z, X] += c * (l2m * ezz + lam * (exx + eyy))
    f["txy"][Y, Z, X] += c * med.mu_xy[Y, Z, X] * (exy + eyx)
    f["txz"][Y, Z, X] += c * med.mu_xz[Y, Z, X] * (exz + ezx)
    f["tyz"][Y, Z, X] += c * med.mu_yz[Y, Z, X] * (eyz + ezy)


def velocity_sweep(F: Fields, med: Medium, cp: Cpml, c: float, y0=None, y1=None):
    """kernel_Vx_Vy_Vz, M:133-188, over interior y in [y0, y1)."""
    ny, nz, nx = F.shape
    y0 = H_STENCIL if y0 is None else y0
    y1 = ny - H_STENCIL if y1 is None else y1
    if y1 <= y0:
        return
    Y, Ym, Yp = slice(y0, y1), slice(y0 - 1, y1 - 1), slice(y0 + 1, y1 + 1)
    Z, Zm, Zp = _I(nz), _S(nz, -1), _S(nz, 1)
    X, Xm, Xp = _I(nx), _S(nx, -1), _S(nx, 1)
    f, P = F.f, F.psi
    txx, tyy, tzz, txy, txz, tyz = (f[k] for k in ("txx", "tyy", "tzz", "txy", "txz", "tyz"))
    ax, bx, ay, by, az, bz = _cp_views(cp, Y, Z, X)
    # M:161-170
    dtxx_dx = txx[Y, Z, Xp] - txx[Y, Z, X]
    dtzz_dz = tzz[Y, Zp, X] - tzz[Y, Z, X]
    dtyy_dy = tyy[Yp, Z, X] - tyy[Y, Z, X]
    txyc, txzc, tyzc = txy[Y, Z, X], txz[Y, Z, X], tyz[Y, Z, X]
    dtxy_dy = txyc - txy[Ym, Z, X]
    dtxy_dx = txyc - txy[Y, Z, Xm]
    dtxz_dz = txzc - txz[Y, Zm, X]
    dtxz_dx = txzc - txz[Y, Z, Xm]
    dtyz_dz = tyzc - tyz[Y, Zm, X]
    dtyz_dy = tyzc - tyz[Ym, Z, X]
    # M:172-182
    fxx = dtxx_dx + _psi_update(P["txx_x"], Y, Z, X, ax, bx, dtxx_dx)
    fyx = dtxy_dx + _psi_update(P["txy_x"], Y, Z, X, ax, bx, dtxy_dx)
    fzx = dtxz_dx + _psi_update(P["txz_x"], Y, Z, X, ax, bx, dtxz_dx)
    fyy = dtyy_dy + _psi_update(P["tyy_y"], Y, Z, X, ay, by, dtyy_dy)
    fxy = dtxy_dy + _psi_update(P["txy_y"], Y, Z, X, ay, by, dtxy_dy)
    fzy = dtyz_dy + _psi_update(P["tyz_y"], Y, Z, X, ay, by, dtyz_dy)
    fzz = dtzz_dz + _psi_update(P["tzz_z"], Y, Z, X, az, bz, dtzz_dz)
    fxz = dtxz_dz + _psi_update(P["txz_z"], Y, Z, X, az, bz, dtxz_dz)
    fyz = dtyz_dz + _psi_update(P["tyz_z"], Y, Z, X, az, bz, dtyz_dz)
    # M:184-186
    f["vx"][Y, Z, X] += c * med.bbx[Y, Z, X] * (fxx + fxy + fxz)
    f["vy"][Y, Z, X] += c * med.bby[Y, Z, X] * (fyx + fyy + fyz)
    f["vz"][Y, Z, X] += c * med.bbz[Y, Z, X] * (fzx + fzy + fzz)


def add_source(F: Fields, pos, m6, amp):
    """addSource, M:1728-1734: τ_ij[Sy,Sz,Sx] += M_ij·amp for all six τ (Q9)."""
    sx, sy, sz = pos
    for k, m in zip(("txx", "tyy", "tzz", "txy", "txz", "tyz"), m6):
        F.f[k][sy, sz, sx] += m * amp


def record(F: Fields, rec):
    """kernel_get_Station, M:324-335.  ``rec`` is (nrec, 3) integer (x, y, z)."""
    rec = np.asarray(rec)
    x, y, z = rec[:, 0], rec[:, 1], rec[:, 2]
    return np.stack([F.f[k][y, z, x] for k in ("vx", "vy", "vz")])


class _Pool:
    """y-slab thread pool for the sweeps (numpy releases the GIL in ufuncs)."""

    def __init__(self, threads):
        self.threads = max(1, int(threads))
        self.ex = ThreadPoolExecutor(self.threads) if self.threads > 1 else None

    def run(self, fn, F, *args):
        ny = F.shape[0]
        lo, hi = H_STENCIL, ny - H_STENCIL
        if self.ex is None:
            fn(F, *args)
            return
        cuts = np.linspace(lo, hi, self.threads + 1).astype(int)
        futs = [self.ex.submit(fn, F, *args, int(cuts[i]), int(cuts[i + 1]))
                for i in range(self.threads) if cuts[i + 1] > cuts[i]]
        for fu in futs:
            fu.result()

    def close(self):
        if self.ex is not None:
            self.ex.shutdown()


def forward(med, cp, c, src_pos, m6, wavelet, rec, nt, threads=1, F=None, history=None,
            borders=None, W=None):
    """forwardFields, M:1812-1848: per step τ → source → v → sample (Q8).

    ``history`` (a list) receives a copy of the pre-step state for every step
    plus the final state (used by the exact adjoint).  ``borders`` (a dict)
    receives the reference boundary save of v (M:1838-1840, 378-419)."""
    F = Fields(med.shape) if F is None else F
    nrec = len(rec)
    g = np.zeros((3, nt, nrec))
    pool = _Pool(threads)
    if borders is not None:
        _borders_alloc(borders, med.shape, nt)
    try:
        for it in range(nt):
            if history is not None:
                history.append(F.copy())
            pool.run(stress_sweep, F, med, cp, c)
            add_source(F, src_pos, m6, wavelet[it])
            pool.run(velocity_sweep, F, med, cp, c)
            g[:, it, :] = record(F, rec)
            if borders is not None:
                get_border(F, borders, it, W)
        if history is not None:
            history.append(F.copy())
    finally:
        pool.close()
    return g, F


# ----------------------------------------------------------------------------
# reference boundary save/restore, CPML zeroing (M:378-499)
# ----------------------------------------------------------------------------
def _borders_alloc(borders, shape, nt):
    ny, nz, nx = shape
    H = H_STENCIL
    for k in ("vx", "vy", "vz"):
        borders[k] = dict(left=np.zeros((nt, ny, nz, H)), right=np.zeros((nt, ny, nz, H)),
                          top=np.zeros((nt, ny, H, nx)), bottom=np.zeros((nt, ny, H, nx)),
                          front=np.zeros((nt, H, nz, nx)), back=np.zeros((nt, H, nz, nx)))


def _border_slices(shape, W):
    ny, nz, nx = shape
    H = H_STENCIL
    return dict(
        left=(slice(None), slice(None), slice(W, W + H)),
        right=(slice(None), slice(None), slice(nx - W - H, nx - W)),
        top=(slice(None), slice(H, 2 * H), slice(None)),
        bottom=(slice(None), slice(nz - W - H, nz - W), slice(None)),
        front=(slice(W, W + H), slice(None), slice(None)),
        back=(slice(ny - W - H, ny - W), slice(None), slice(None)),
    )


def get_border(F, borders, it, W):
    """kernel_get_border ×3 (vx, vy, vz), M:378-419 / M:1524-1562."""
    sl = _border_slices(F.shape, W)
    for k in ("vx", "vy", "vz"):
        for face, s in sl.items():
            borders[k][face][it] = F.f[k][s]


def set_border(F, borders, it, W):
    """kernel_set_border ×3, M:422-462 / M:1692-1726."""
    sl = _border_slices(F.shape, W)
    for k in ("vx", "vy", "vz"):
        for face, s in sl.items():
            F.f[k][s] = borders[k][face][it]


def zero_cpml_zone(F, W):
    """Fields.setZeroCPMLzone, M:1007-1071 / kernel M:466-499: zero x, y and
    bottom CPML zones (not the top) of v and τ."""
    ny, nz, nx = F.shape
    for k in FIELDS:
        a = F.f[k]
        a[:, :, :W] = 0.0
        a[:, :, nx - W:] = 0.0
        a[:, nz - W:, :] = 0.0
        a[:W, :, :] = 0.0
        a[ny - W:, :, :] = 0.0


# ----------------------------------------------------------------------------
# reference-style adjoint (Q13, Q14, Q16, Q22)
# ----------------------------------------------------------------------------
def add_adjoint_source(F, rec, res, it, nt):
    """kernel_add_adjoint_source, M:310-322: v[rec] += res[nt-it-1] (time reversed)."""
    rec = np.asarray(rec)
    x, y, z = rec[:, 0], rec[:, 1], rec[:, 2]
    for ci, k in enumerate(("vx", "vy", "vz")):
        np.add.at(F.f[k], (y, z, x), res[ci, nt - it - 1])


def image_mt(B, med_f32_lam, med_f32_mu, src_pos, w):
    """kernel_Gradients_Mij, M:337-372 (compliance of the adjoint stress at the
    source, times the reversed wavelet sample ``w``).  Returns 6 values."""
    sx, sy, sz = src_pos
    lam = float(med_f32_lam[sy, sz, sx]); mu = med_f32_mu
    Mu = float(mu[sy, sz, sx]); Mi = float(mu[sy, sz, sx + 1]); Mk = float(mu[sy, sz + 1, sx])
    Mj = float(mu[sy + 1, sz, sx]); Mik = float(mu[sy, sz + 1, sx + 1])
    Mjk = float(mu[sy + 1, sz + 1, sx]); Mij = float(mu[sy + 1, sz, sx + 1])
    mxz = 1 / (0.25 / Mu + 0.25 / Mi + 0.25 / Mk + 0.25 / Mik)
    mxy = 1 / (0.25 / Mu + 0.25 / Mi + 0.25 / Mj + 0.25 / Mij)
    myz = 1 / (0.25 / Mu + 0.25 / Mj + 0.25 / Mk + 0.25 / Mjk)
    t = {k: float(B.f[k][sy, sz, sx]) for k in ("txx", "tyy", "tzz", "txy", "txz", "tyz")}
    comp = 1.0 / (2.0 * Mu * (3.0 * lam + 2.0 * Mu))
    return np.array([
        w * comp * (2.0 * (lam + Mu) * t["txx"] - lam * t["tyy"] - lam * t["tzz"]),
        w * comp * (2.0 * (lam + Mu) * t["tyy"] - lam * t["txx"] - lam * t["tzz"]),
        w * comp * (2.0 * (lam + Mu) * t["tzz"] - lam * t["txx"] - lam * t["tyy"]),
        w * t["txy"] / mxy, w * t["txz"] / mxz, w * t["tyz"] / myz])


def ref_backward_mt(med, cp, c, src_pos, wavelet, rec, nt, res, threads=1):
    """backwardFields (mode 0), M:2056-2082: the forward operator with +dt run on a
    zeroed adjoint state, residual injected time-reversed into v, MT image
    accumulated every step (nt steps, Q23).  ``res`` is (3, nt, nrec)."""
    B = Fields(med.shape)
    g = np.zeros(6)
    pool = _Pool(threads)
    try:
        for it in range(nt):
            pool.run(stress_sweep, B, med, cp, c)
            pool.run(velocity_sweep, B, med, cp, c)
            add_adjoint_source(B, rec, res, it, nt)
            g += image_mt(B, med.lam, med.mu, src_pos, float(wavelet[nt - it - 1]))
    finally:
        pool.close()
    return g


def _image_ref(B, Fw, Vpres, med, dt, dx, dy, dz):
    """kernel_Gradients_Rho / _Lambda / _Mu, M:191-294 (interior only)."""
    ny, nz, nx = Fw.shape
    Y, Z, X = _I(ny), _I(nz), _I(nx)
    Ym, Zm, Xm = _S(ny, -1), _S(nz, -1), _S(nx, -1)
    Ymm, Zmm, Xmm = _S(ny, -2), _S(nz, -2), _S(nx, -2)
    Yp, Zp, Xp = _S(ny, 1), _S(nz, 1), _S(nx, 1)
    b, f = B.f, Fw.f
    grho = np.zeros(Fw.shape); glam = np.zeros(Fw.shape); gmu = np.zeros(Fw.shape)
    # M:206
    grho[Y, Z, X] = -((b["vx"][Y, Z, X] * (Vpres["vx"][Y, Z, X] - f["vx"][Y, Z, X]) / dt)
                      + (b["vy"][Y, Z, X] * (Vpres["vy"][Y, Z, X] - f["vy"][Y, Z, X]) / dt)
                      + (b["vz"][Y, Z, X] * (Vpres["vz"][Y, Z, X] - f["vz"][Y, Z, X]) / dt))
    # M:230-238 (fields are v_present)
    vx, vy, vz = Vpres["vx"], Vpres["vy"], Vpres["vz"]
    lam = med.lam[Y, Z, X]; mu = med.mu[Y, Z, X]
    dvx = 0.5 * (vx[Y, Z, X] - vx[Y, Z, Xm]) / dx + 0.5 * (vx[Y, Z, Xm] - vx[Y, Z, Xmm]) / dx
    dvy = 0.5 * (vy[Y, Z, X] - vy[Ym, Z, X]) / dy + 0.5 * (vy[Ym, Z, X] - vy[Ymm, Z, X]) / dy
    dvz = 0.5 * (vz[Y, Z, X] - vz[Y, Zm, X]) / dz + 0.5 * (vz[Y, Zm, X] - vz[Y, Zmm, X]) / dz
    txx, tyy, tzz = b["txx"][Y, Z, X], b["tyy"][Y, Z, X], b["tzz"][Y, Z, X]
    glam[Y, Z, X] = (txx + tyy + tzz) * (dvx + dvy + dvz) / (3.0 * lam + 2 * mu)
    # M:269-291
    ex = vx[Y, Z, X] - vx[Y, Z, Xm]
    ey = vy[Y, Z, X] - vy[Ym, Z, X]
    ez = vz[Y, Z, X] - vz[Y, Zm, X]
    dvx_dz = vx[Y, Zp, X] - vx[Y, Z, X]
    dvx_dy = vx[Yp, Z, X] - vx[Y, Z, X]
    dvy_dx = vy[Y, Z, Xp] - vy[Y, Z, X]
    dvy_dz = vy[Y, Zp, X] - vy[Y, Z, X]
    dvz_dx = vz[Y, Z, Xp] - vz[Y, Z, X]
    dvz_dy = vz[Yp, Z, X] - vz[Y, Z, X]
    k = 1.0 / (mu * (3 * lam + 2 * mu))
    gmu[Y, Z, X] = (k * (2 * (lam + mu) * txx - lam * tyy - lam * tzz) * ex
                    + k * (2 * (lam + mu) * tyy - lam * txx - lam * tzz) * ey
                    + k * (2 * (lam + mu) * tzz - lam * txx - lam * tyy) * ez
                    + (1.0 / med.mu_xy[Y, Z, X]) * b["txy"][Y, Z, X] * (dvy_dx + dvx_dy)
                    + (1.0 / med.mu_xz[Y, Z, X]) * b["txz"][Y, Z, X] * (dvx_dz + dvz_dx)
                    + (1.0 / med.mu_yz[Y, Z, X]) * b["tyz"][Y, Z, X] * (dvz_dy + dvy_dz))
    return glam, gmu, grho


def ref_backward_remake(med, cp, c, src_pos, m6, wavelet, rec, nt, res, Ffinal, borders, W,
                        dt, dx, dy=None, dz=None, threads=1):
    """backwardFields_FowardFieldsRemake (mode 1), M:1929-2054.

    ``Ffinal`` is the forward state after the last step (ψ included: it is not
    reset, Q22) and is consumed in place.  Returns (gλ, gμ, gρ) accumulated over
    nt-1 iterations (Q23)."""
    dy = dx if dy is None else dy
    dz = dx if dz is None else dz
    B = Fields(med.shape)
    Fw = Ffinal
    glam = np.zeros(med.shape); gmu = np.zeros(med.shape); grho = np.zeros(med.shape)
    pool = _Pool(threads)
    try:
        for it in range(nt - 1):
            pool.run(stress_sweep, B, med, cp, c)
            pool.run(velocity_sweep, B, med, cp, c)
            add_adjoint_source(B, rec, res, it, nt)
            Vpres = {k: Fw.f[k].copy() for k in ("vx", "vy", "vz")}      # kernel_DTD_copy
            pool.run(velocity_sweep, Fw, med, cp, -c)                    # flagMode = 1
            set_border(Fw, borders, nt - it - 2, W)
            zero_cpml_zone(Fw, W)
            pool.run(stress_sweep, Fw, med, cp, -c)
            add_source(Fw, src_pos, m6, -float(wavelet[nt - it - 1]))
            gl, gm, gr = _image_ref(B, Fw, Vpres, med, dt, dx, dy, dz)
            glam += gl; gmu += gm; grho += gr
    finally:
        pool.close()
    return glam, gmu, grho


# ----------------------------------------------------------------------------
# exact discrete adjoint (transpose of the forward recurrence)
# ----------------------------------------------------------------------------
def _full(shape):
    return np.zeros(shape)


def exact_adjoint(med, cp, c, src_pos, wavelet, rec, nt, resid, history):
    """Reverse-mode derivative of the forward recurrence above.

    ``resid`` (3, nt, nrec) is ∂J/∂gathers; ``history`` is the list written by
    ``forward(..., history=[])`` (pre-step states S^0..S^nt).  Returns a dict of
    ∂J/∂λ, ∂J/∂μ, ∂J/∂ρ (volumes) and ∂J/∂M (6), for fixed wavelet and CPML.

    Per reverse step n:  R1 v̄[rec] += r[n];  R2 transpose of the velocity sweep
    (ρ accumulators, τ̄ += D_τᵀ h̄);  R3 M̄ += w[n] τ̄(src);  R4 transpose of the
    stress sweep (λ, μ accumulators, v̄ += D_vᵀ ḡ).  ψ adjoints are carried in
    C (C ← b(C + f̄), h̄ = f̄ + a(C + f̄))."""
    shape = med.shape
    ny, nz, nx = shape
    Y, Z, X = _I(ny), _I(nz), _I(nx)
    Ym, Zm, Xm = _S(ny, -1), _S(nz, -1), _S(nx, -1)
    Yp, Zp, Xp = _S(ny, 1), _S(nz, 1), _S(nx, 1)
    rec = np.asarray(rec)
    ax, bx, ay, by, az, bz = _cp_views(cp, Y, Z, X)
    cpk = {}
    for k in PSI_V + PSI_T:
        axis = k[-1]
        cpk[k] = {"x": (ax, bx), "y": (ay, by), "z": (az, bz)}[axis]
    vb = {k: _full(shape) for k in ("vx", "vy", "vz")}
    tb = {k: _full(shape) for k in ("txx", "tyy", "tzz", "txy", "txz", "tyz")}
    Cv = {k: _full(shape)[Y, Z, X].copy() for k in PSI_V}
    Ct = {k: _full(shape)[Y, Z, X].copy() for k in PSI_T}
    G = {k: _full(shape)[Y, Z, X].copy() for k in ("lam", "mu", "xy", "xz", "yz", "rx", "ry", "rz")}
    gm6 = np.zeros(6)
    lam = med.lam[Y, Z, X]; mu = med.mu[Y, Z, X]; l2m = lam + 2.0 * mu
    sx, sy, sz = src_pos

    def carry(Cd, k, fbar):
        a, b = cpk[k]
        Psi = Cd[k] + fbar
        out = fbar + a * Psi
        Cd[k] = b * Psi
        return out

    for n in range(nt - 1, -1, -1):
        Sn, Sn1 = history[n], history[n + 1]
        # R1
        for ci, k in enumerate(("vx", "vy", "vz")):
            np.add.at(vb[k], (rec[:, 1], rec[:, 2], rec[:, 0]), resid[ci, n])
        # R2: forward f at step n from τ^{n+1} (= τ'') and ψτ^n
        t = Sn1.f
        h = dict(
            txx_x=t["txx"][Y, Z, Xp] - t["txx"][Y, Z, X], tyy_y=t["tyy"][Yp, Z, X] - t["tyy"][Y, Z, X],
            tzz_z=t["tzz"][Y, Zp, X] - t["tzz"][Y, Z, X], txy_y=t["txy"][Y, Z, X] - t["txy"][Ym, Z, X],
            txy_x=t["txy"][Y, Z, X] - t["txy"][Y, Z, Xm], txz_z=t["txz"][Y, Z, X] - t["txz"][Y, Zm, X],
            txz_x=t["txz"][Y, Z, X] - t["txz"][Y, Z, Xm], tyz_z=t["tyz"][Y, Z, X] - t["tyz"][Y, Zm, X],
            tyz_y=t["tyz"][Y, Z, X] - t["tyz"][Ym, Z, X])
        fv = {}
        for k in PSI_T:
            a, b = cpk[k]
            fv[k] = h[k] + (b * Sn.psi[k][Y, Z, X] + a * h[k])
        Fx = fv["txx_x"] + fv["txy_y"] + fv["txz_z"]
        Fy = fv["txy_x"] + fv["tyy_y"] + fv["tyz_z"]
        Fz = fv["txz_x"] + fv["tyz_y"] + fv["tzz_z"]
        G["rx"] += vb["vx"][Y, Z, X] * Fx
        G["ry"] += vb["vy"][Y, Z, X] * Fy
        G["rz"] += vb["vz"][Y, Z, X] * Fz
        ux = c * med.bbx[Y, Z, X] * vb["vx"][Y, Z, X]
        uy = c * med.bby[Y, Z, X] * vb["vy"][Y, Z, X]
        uz = c * med.bbz[Y, Z, X] * vb["vz"][Y, Z, X]
        src_u = dict(txx_x=ux, txy_y=ux, txz_z=ux, txy_x=uy, tyy_y=uy, tyz_z=uy,
                     txz_x=uz, tyz_y=uz, tzz_z=uz)
        hb = {}
        for k in PSI_T:
            full = _full(shape)
            full[Y, Z, X] = carry(Ct, k, src_u[k])
            hb[k] = full
        tb["txx"][Y, Z, X] += hb["txx_x"][Y, Z, Xm] - hb["txx_x"][Y, Z, X]
        tb["tyy"][Y, Z, X] += hb["tyy_y"][Ym, Z, X] - hb["tyy_y"][Y, Z, X]
        tb["tzz"][Y, Z, X] += hb["tzz_z"][Y, Zm, X] - hb["tzz_z"][Y, Z, X]
        tb["txy"][Y, Z, X] += (hb["txy_y"][Y, Z, X] - hb["txy_y"][Yp, Z, X]
                               + hb["txy_x"][Y, Z, X] - hb["txy_x"][Y, Z, Xp])
        tb["txz"][Y, Z, X] += (hb["txz_z"][Y, Z, X] - hb["txz_z"][Y, Zp, X]
                               + hb["txz_x"][Y, Z, X] - hb["txz_x"][Y, Z, Xp])
        tb["tyz"][Y, Z, X] += (hb["tyz_z"][Y, Z, X] - hb["tyz_z"][Y, Zp, X]
                               + hb["tyz_y"][Y, Z, X] - hb["tyz_y"][Yp, Z, X])
        # R3
        w = float(wavelet[n])
        for i, k in enumerate(("txx", "tyy", "tzz", "txy", "txz", "tyz")):
            gm6[i] += w * tb[k][sy, sz, sx]
        # R4: forward e at step n from v^n and ψv^n
        v = Sn.f
        g = dict(
            vx_x=v["vx"][Y, Z, X] - v["vx"][Y, Z, Xm], vy_y=v["vy"][Y, Z, X] - v["vy"][Ym, Z, X],
            vz_z=v["vz"][Y, Z, X] - v["vz"][Y, Zm, X], vx_z=v["vx"][Y, Zp, X] - v["vx"][Y, Z, X],
            vx_y=v["vx"][Yp, Z, X] - v["vx"][Y, Z, X], vy_x=v["vy"][Y, Z, Xp] - v["vy"][Y, Z, X],
            vy_z=v["vy"][Y, Zp, X] - v["vy"][Y, Z, X], vz_x=v["vz"][Y, Z, Xp] - v["vz"][Y, Z, X],
            vz_y=v["vz"][Yp, Z, X] - v["vz"][Y, Z, X])
        e = {}
        for k in PSI_V:
            a, b = cpk[k]
            e[k] = g[k] + (b * Sn.psi[k][Y, Z, X] + a * g[k])
        txx, tyy, tzz = tb["txx"][Y, Z, X], tb["tyy"][Y, Z, X], tb["tzz"][Y, Z, X]
        txy, txz, tyz = tb["txy"][Y, Z, X], tb["txz"][Y, Z, X], tb["tyz"][Y, Z, X]
        G["lam"] += (txx + tyy + tzz) * (e["vx_x"] + e["vy_y"] + e["vz_z"])
        G["mu"] += 2.0 * (txx * e["vx_x"] + tyy * e["vy_y"] + tzz * e["vz_z"])
        G["xy"] += txy * (e["vx_y"] + e["vy_x"])
        G["xz"] += txz * (e["vx_z"] + e["vz_x"])
        G["yz"] += tyz * (e["vy_z"] + e["vz_y"])
        eb_xx = c * (l2m * txx + lam * (tyy + tzz))
        eb_yy = c * (l2m * tyy + lam * (txx + tzz))
        eb_zz = c * (l2m * tzz + lam * (txx + tyy))
        eb_xy = c * med.mu_xy[Y, Z, X] * txy
        eb_xz = c * med.mu_xz[Y, Z, X] * txz
        eb_yz = c * med.mu_yz[Y, Z, X] * tyz
        src_e = dict(vx_x=eb_xx, vy_y=eb_yy, vz_z=eb_zz, vx_y=eb_xy, vy_x=eb_xy,
                     vx_z=eb_xz, vz_x=eb_xz, vy_z=eb_yz, vz_y=eb_yz)
        gb = {}
        for k in PSI_V:
            full = _full(shape)
            full[Y, Z, X] = carry(Cv, k, src_e[k])
            gb[k] = full
        vb["vx"][Y, Z, X] += (gb["vx_x"][Y, Z, X] - gb["vx_x"][Y, Z, Xp]
                              + gb["vx_z"][Y, Zm, X] - gb["vx_z"][Y, Z, X]
                              + gb["vx_y"][Ym, Z, X] - gb["vx_y"][Y, Z, X])
        vb["vy"][Y, Z, X] += (gb["vy_y"][Y, Z, X] - gb["vy_y"][Yp, Z, X]
                              + gb["vy_x"][Y, Z, Xm] - gb["vy_x"][Y, Z, X]
                              + gb["vy_z"][Y, Zm, X] - gb["vy_z"][Y, Z, X])
        vb["vz"][Y, Z, X] += (gb["vz_z"][Y, Z, X] - gb["vz_z"][Y, Zp, X]
                              + gb["vz_x"][Y, Z, Xm] - gb["vz_x"][Y, Z, X]
                              + gb["vz_y"][Ym, Z, X] - gb["vz_y"][Y, Z, X])
    return combine_gradients(med, c, G, gm6)


def combine_gradients(med, c, G, gm6):
    """Turn per-cell accumulators into ∂J/∂λ, ∂J/∂μ, ∂J/∂ρ.

    μ̄_xy(c') = 4/Σ 1/μ over {c', c'+x, c'+y, c'+x+y}: ∂μ̄/∂μ_k = μ̄²/(4μ_k²).
    b̄_x(c') = 2/(ρ(c')+ρ(c'+x)): ∂b̄/∂ρ = -b̄²/2 for both cells."""
    shape = med.shape
    ny, nz, nx = shape
    Y, Z, X = _I(ny), _I(nz), _I(nx)

    def pad(a):
        f = np.zeros(shape)
        f[Y, Z, X] = a
        return f

    glam = c * pad(G["lam"])
    gmu = c * pad(G["mu"])
    with np.errstate(divide="ignore", invalid="ignore"):
        for key, mbar, (o1, o2) in (("xy", med.mu_xy, ((0, 0, 1), (1, 0, 0))),
                                    ("xz", med.mu_xz, ((0, 0, 1), (0, 1, 0))),
                                    ("yz", med.mu_yz, ((1, 0, 0), (0, 1, 0)))):
            w = c * pad(G[key]) * mbar * mbar / 4.0
            acc = np.zeros(shape)
            for (dy, dz, dx) in ((0, 0, 0), o1, o2, tuple(a + b for a, b in zip(o1, o2))):
                acc[dy:, dz:, dx:] += w[:ny - dy, :nz - dz, :nx - dx]
            gmu += np.where(med.mu != 0, acc / (med.mu * med.mu), 0.0)
    grho = np.zeros(shape)
    for key, bb, (dy, dz, dx) in (("rx", med.bbx, (0, 0, 1)), ("ry", med.bby, (1, 0, 0)),
                                  ("rz", med.bbz, (0, 1, 0))):
        w = -0.5 * c * pad(G[key]) * bb * bb
        grho += w
        grho[dy:, dz:, dx:] += w[:ny - dy, :nz - dz, :nx - dx]
    return dict(lam=glam, mu=gmu, rho=grho, m6=gm6)


def velocity_gradients(vp, vs, rho, g):
    """Chain rule from (λ, μ, ρ) to (Vp, Vs, ρ) for λ = ρVp² - 2ρVs², μ = ρVs²."""
    vp = np.asarray(vp, np.float64); vs = np.asarray(vs, np.float64); rho = np.asarray(rho, np.float64)
    gvp = g["lam"] * 2.0 * rho * vp
    gvs = g["lam"] * (-4.0 * rho * vs) + g["mu"] * 2.0 * rho * vs
    grho = g["lam"] * (vp * vp - 2.0 * vs * vs) + g["mu"] * vs * vs + g["rho"]
    return dict(vp=gvp, vs=gvs, rho=grho, m6=g["m6"])


def misfit(g_mod, g_obs):
    """computeResidual / computeCostFunction, M:2128-2142: r = mod - obs, J = ½‖r‖²."""
    r = np.asarray(g_mod, np.float64) - np.asarray(g_obs, np.float64)
    return r, 0.5 * float(np.sum(r * r))


def xcorr_residual(g_mod, g_obs):
    """computeResidualCrossCorrelation, M:2145-2179, per trace (component c,
    receiver r) of (3, nt, nrec) gathers: with a = ⟨m, o⟩, |m|, |o| over time,

        res = m·a/(|o|·|m|³) - o/(|m|·|o|)

    (the derivative of 1 - a/(|m||o|) with respect to m), where a zero |m|, |o|
    or a is replaced by 1 as the reference does.  Also returns the per-trace
    cost Σ (1 - a/(|m||o|)) with the same guarded norms (raw a), which that
    residual differentiates."""
    m = np.asarray(g_mod, np.float64)
    o = np.asarray(g_obs, np.float64)
    nm = np.sqrt(np.sum(m * m, axis=-2, keepdims=True))
    no = np.sqrt(np.sum(o * o, axis=-2, keepdims=True))
    a = np.sum(m * o, axis=-2, keepdims=True)
    nm = np.where(nm == 0, 1.0, nm)
    no = np.where(no == 0, 1.0, no)
    cost = float(np.sum(1.0 - a / (nm * no)))
    ag = np.where(a == 0, 1.0, a)
    res = m * (ag / (no * nm ** 3)) - o / (nm * no)
    return res, cost


def xcorr_cost(g_mod, g_obs):
    """computeCostFunctionCrossCorrelation, M:2181-2206: per component, one
    minus the normalised correlation of the whole (nt, nrec) gathers (zero norms
    replaced by 1), summed over the three components."""
    m = np.asarray(g_mod, np.float64)
    o = np.asarray(g_obs, np.float64)
    total = 0.0
    for c in range(3):
        nm = np.sqrt(np.sum(m[c] * m[c])) or 1.0
        no = np.sqrt(np.sum(o[c] * o[c])) or 1.0
        total += 1.0 - float(np.sum(m[c] * o[c])) / (nm * no)
    return total


def default_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
