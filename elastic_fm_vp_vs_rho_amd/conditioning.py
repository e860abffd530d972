"""Device-side gradient conditioning and model maps of the FWI driver
(libefm.so: efm_condition_gradient, efm_model_update, efm_model_rho_nafe_drake,
efm_volume_axpy).

Reference: fwi_Vp_Vs_Rho (mainPropElastic3D_MT.py M:2349-2719) smooths the
summed images with scipy.ndimage.gaussian_filter(σ=1) and zeroes the top 10
cells and the CPML+2 zone (M:2479-2504), then updates λ and μ and maps them to
Vp, Vs (M:1200-1202) and ρ (Nafe-Drake, M:1204-1206) on the host.  Here the
volumes stay on the device; torch only owns the memory.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi
from ._capi import EfmSmooth, check
from .propagator import _ptr, _stream


def reference_margins(nx, ny, nz, W, top=10):
    """Face margins zeroed by the reference (M:2483-2504): z < top, and W+2
    cells at x_lo, x_hi, y_lo, y_hi and z_hi.  Order: (x_lo, x_hi, y_lo, y_hi,
    z_lo, z_hi)."""
    m = W + 2
    return (m, m, m, m, top, m)


def _volume(t, name):
    if t.dtype != torch.float32 or not t.is_contiguous() or t.dim() != 3:
        raise ValueError(f"{name} must be a contiguous float32 (ny, nz, nx) volume")
    return t


def condition_gradient(g, sigma=1.0, truncate=4.0, margins=(0, 0, 0, 0, 0, 0), out=None, work=None):
    """mask(gaussian_filter(g, σ, mode='reflect', truncate)) of one device volume."""
    _volume(g, "gradient")
    ny, nz, nx = g.shape
    out = torch.empty_like(g) if out is None else out
    if sigma > 0 and work is None:
        work = torch.empty_like(g)
    p = EfmSmooth(float(sigma), float(truncate), (ctypes.c_int * 6)(*[int(v) for v in margins]))
    check(_capi.lib().efm_condition_gradient(ctypes.byref(p), nx, ny, nz, _ptr(g), _ptr(out),
                                             _ptr(work) if work is not None else None, _stream()),
          "efm_condition_gradient")
    return out


def model_update(lam0, mu0, rho_in, dlam=None, alpha_lam=0.0, dmu=None, alpha_mu=0.0,
                 rho_keep=None, top=0, with_rho=True):
    """(λ, μ, Vp, Vs, ρ) after λ = λ0 + α_λ·dλ, μ = μ0 + α_μ·dμ, Vp/Vs from
    (λ, μ, rho_in), ρ = Nafe-Drake(Vp) except z < top where rho_keep is kept
    (ρ is None when with_rho is False)."""
    for t, n in ((lam0, "lam0"), (mu0, "mu0"), (rho_in, "rho_in")):
        _volume(t, n)
    ny, nz, nx = lam0.shape
    lam, mu, vp, vs = (torch.empty_like(lam0) for _ in range(4))
    rho = torch.empty_like(lam0) if with_rho else None
    check(_capi.lib().efm_model_update(_ptr(lam0), _ptr(dlam), float(alpha_lam), _ptr(mu0), _ptr(dmu),
                                       float(alpha_mu), _ptr(rho_in), _ptr(rho_keep), int(top), nx, ny,
                                       nz, _ptr(lam), _ptr(mu), _ptr(vp), _ptr(vs), _ptr(rho), _stream()),
          "efm_model_update")
    return lam, mu, vp, vs, rho


def rho_nafe_drake(vp):
    rho = torch.empty_like(vp)
    check(_capi.lib().efm_model_rho_nafe_drake(_ptr(vp), _ptr(rho), vp.numel(), _stream()),
          "efm_model_rho_nafe_drake")
    return rho


def axpy(a, x, y):
    """y += a·x in place (float32)."""
    if x.shape != y.shape or x.dtype != torch.float32 or y.dtype != torch.float32:
        raise ValueError("axpy needs two float32 arrays of one shape")
    check(_capi.lib().efm_volume_axpy(float(a), _ptr(x.contiguous()), _ptr(y), y.numel(), _stream()),
          "efm_volume_axpy")
    return y
