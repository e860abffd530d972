"""efm_gradient_static (the allocation-free whole gradient) against the path
with caller workspaces (efm_derived_medium + efm_adjoint + efm_gradient_combine,
through Propagator.backprop): identical bits, on two calls in a row, the second
on a smaller grid, so that what the first call left in the library's static
scratch must not leak into the second.  Collected last: it shares the
library's static scratch with nothing else, and passes alone and after every
other test."""
import ctypes

import numpy as np
import pytest
import torch

from tests.gpu_helpers import propagator_of
from tests.problems import tiny

pytestmark = pytest.mark.gpu


def _static(L):
    return L.efm_gradient_static


def _case(p, obs_scale):
    from elastic_fm_vp_vs_rho_amd.propagator import _ptr, _stream
    from elastic_fm_vp_vs_rho_amd import _capi
    prop = propagator_of(p)
    tape = prop.record(strategy="store")
    obs = (obs_scale * tape["gathers"]).contiguous()
    resid, _ = prop.misfit(tape["gathers"], obs)
    glam, gmu, grho, gm6 = prop.backprop(tape, resid)
    # the same gradient through the static-scratch entry point, with a dirty adjoint state
    adj = prop.new_wavefield()
    adj.fields.fill_(3.0)
    adj.psi_x.fill_(-1.0); adj.psi_y.fill_(2.0); adj.psi_z.fill_(5.0)
    out = [torch.full(prop.grid.shape, 7.0, device="cuda") for _ in range(3)]
    g6 = torch.full((prop.shots.S, 6), 9.0, dtype=torch.float64, device="cuda")
    d, m, sh = prop.dims(), prop.medium(), prop.shots.struct()
    fn = _static(_capi.lib())
    print(f"efm_gradient_static: grid {p.nx}x{p.ny}x{p.nz}, nt {p.nt}", flush=True)
    rc = fn(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh), prop._array(tape["states"]),
            ctypes.byref(adj.struct()), _ptr(resid), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), _ptr(g6),
            _stream())
    assert rc == 0, _capi.last_error()
    torch.cuda.synchronize()
    for a, b, name in zip(out, (glam, gmu, grho), ("lam", "mu", "rho")):
        assert torch.equal(a, b), name
    assert torch.equal(g6, gm6)
    assert float(glam.abs().max()) > 0 and float(gm6.abs().max()) > 0


def test_static_gradient_matches_workspace_path_twice(gpu_lib):
    _case(tiny(n=(30, 26, 22), nt=70, W=4, seed=41, nrec=7), 0.7)
    _case(tiny(n=(20, 18, 16), nt=45, W=3, seed=42, nrec=5, free_top=True), 1.3)


def test_static_gradient_refuses_oversized_batches(gpu_lib):
    from elastic_fm_vp_vs_rho_amd._capi import EfmDims, EfmMedium, EfmShots, EfmState
    from elastic_fm_vp_vs_rho_amd import _capi
    L = _capi.lib()
    fn = _static(L)
    d = EfmDims(160, 160, 160, 1, 10, 10, 10, 10, 0, 10, 1, 0, 0.1)     # 4.1e6 cells: over 64 MiB
    st = EfmState(*([1] * 12))
    fwd = (ctypes.POINTER(EfmState) * 2)(ctypes.pointer(st), ctypes.pointer(st))
    rc = fn(ctypes.byref(d), ctypes.byref(EfmMedium(1, 1, 1)), ctypes.byref(EfmShots(*([1] * 12))), fwd,
            ctypes.byref(st), None, 1, 1, 1, 1, None)
    assert rc == -1 and "scratch" in _capi.last_error()
