"""Host-side model maps, CPML profiles and source wavelets (numpy).

These are the setup computations the reference performs on the host once per
model / frequency; they are restated here in vectorised form.  Device-side
Lamé maps for the autograd path are efm_model_lame / efm_model_lame_adjoint.
"""
from __future__ import annotations

import numpy as np

REF_DT = 0.08551979923718353   # Wavelet.__init__ hard-coded sample interval (M:1254)


def lame_from_velocity(vp, vs, rho):
    """Model.compute_Lambda_Mu (M:1196-1198): μ = Vs·Vs·ρ, λ = Vp·Vp·ρ - 2μ (float32)."""
    vp = np.asarray(vp, np.float32); vs = np.asarray(vs, np.float32)
    rho = np.asarray(rho, np.float32)
    mu = vs * vs * rho
    return vp * vp * rho - np.float32(2.0) * mu, mu


def velocity_from_lame(lam, mu, rho):
    """Model.compute_Vp_Vs (M:1200-1202)."""
    return np.sqrt((lam + np.float32(2.0) * mu) / rho), np.sqrt(mu / rho)


def rho_nafe_drake(vp):
    """Model.compute_Rho (M:1204-1206): Nafe-Drake polynomial in km/s (the
    reference's quadratic coefficient is 0.4712)."""
    v = np.asarray(vp) / 1e3
    c = (1.6612, -0.4712, 0.0671, -0.0043, 0.000106)
    return (c[0] * v + c[1] * v ** 2 + c[2] * v ** 3 + c[3] * v ** 4 + c[4] * v ** 5) * 1e3


def cpml_profile(frequency, dt, di, W, vmax, N):
    """CPML.CPML_i (M:1340-1372): float32 (a, b) of length N.

    d(F) = d0·vmax·(F/L)², d0 = -3 ln R / (2L), R = 1e-6, L = W·di;
    α(F) = π f (L-F)/L;  b = exp(-(d+α)dt);  a = d/(d+α)·(b-1);
    F = (W-1-i)·di on the low side and (i-(N-W))·di on the high side."""
    a = np.zeros(N, np.float32)
    b = np.zeros(N, np.float32)
    if W <= 0:
        return a, b
    L = W * di
    d0 = -3 * np.log(1e-6) / (2.0 * L)
    i = np.arange(N)
    lo, hi = i < W, i >= N - W
    F = np.where(lo, (W - 1 - i) * di, np.where(hi, (i - (N - W)) * di, 0.0)).astype(np.float64)
    d = d0 * vmax * (F / L) * (F / L)
    alpha = np.pi * frequency * ((L - F) / L)
    bb = np.exp(-(d + alpha) * dt)
    aa = (d / (d + alpha)) * (bb.astype(np.float32).astype(np.float64) - 1.0)
    m = lo | hi
    b[m] = bb[m]
    a[m] = aa[m]
    return a, b


def cpml_coefficients(W, frequency, dt, dx, vmax, nx, ny, nz, free_top=True, dy=None, dz=None):
    """CPML.__init__ (M:1298-1307); free_top zeroes the z-min profile (M:1305-1306)."""
    ax, bx = cpml_profile(frequency, dt, dx, W, vmax, nx)
    ay, by = cpml_profile(frequency, dt, dx if dy is None else dy, W, vmax, ny)
    az, bz = cpml_profile(frequency, dt, dx if dz is None else dz, W, vmax, nz)
    if free_top:
        az[:W] = 0.0
        bz[:W] = 0.0
    return dict(ax=ax, bx=bx, ay=ay, by=by, az=az, bz=bz)


def ricker(fq, dt, nt, fq_exp=None):
    """Wavelet.ricker (M:1264-1272): (1 - 2(πf(t-t0)dt)²)·exp(-(πf'(t-t0)dt)²),
    t0 = √2/f/dt samples; f' is the module-global fq of the reference (M:1270),
    equal to f unless given."""
    fe = fq if fq_exp is None else fq_exp
    t0 = (np.sqrt(2) / fq) / dt
    tau = (np.arange(nt) - t0) * dt
    return ((1 - 2 * (np.pi * fq * tau) ** 2) * np.exp(-(np.pi * fe * tau) ** 2)).astype(np.float32)


def gaussian(fq, dt, nt):
    """Wavelet.gaussian (M:1274-1281)."""
    t0 = (1.0 / fq) / dt
    tau = (np.arange(nt) - t0) * dt
    return np.exp(-2 * np.pi ** 2 * fq ** 2 * tau ** 2).astype(np.float32)
