"""The cross-correlation misfit (M:2145-2206) and the reference's mode-1
workflow around the path (starting model M:3073-3082, frequency continuation
M:3134-3203) on the device, against the FP64 oracle and scipy."""
import os

import numpy as np
import pytest
import torch
from scipy.ndimage import gaussian_filter

from elastic_fm_vp_vs_rho_amd import driver as D
from oracle import elastic_fp64 as O
from tests.problems import rel_l2, run_forward, tiny
from tests.test_gpu_driver import _oracle, _setup

pytestmark = pytest.mark.gpu


def test_xcorr_residual_kernel_against_oracle(gpu_lib):
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, Propagator, ShotBatch
    rng = np.random.default_rng(3)
    S, nt, nrec = 3, 57, 7
    mod = rng.standard_normal((S, 3, nt, nrec)).astype(np.float32)
    obs = rng.standard_normal((S, 3, nt, nrec)).astype(np.float32)
    mod[1, 2, :, 4] = 0.0                                     # a silent modelled trace (guards)
    obs[2, 0, :, 1] = 0.0
    obs[0, 1, :, 3] = 2.5 * mod[0, 1, :, 3]                   # a perfectly correlated one
    # a propagator only to reach the misfit entry point
    g = Grid(8, 8, 8, 10.0, 1e-3, (0,) * 6)
    z = torch.ones(g.shape, device="cuda")
    sh = ShotBatch(g, [(4, 4, 4)], [np.zeros(6)], [np.zeros(nt)], [nt], [[(4, 4, 4)]],
                   {k: np.zeros(8, np.float32) for k in ("ax", "bx", "ay", "by", "az", "bz")})
    prop = Propagator(g, z, z, z, sh)
    r, J = prop.misfit(torch.as_tensor(mod, device="cuda"), torch.as_tensor(obs, device="cuda"), "xcorr")
    Jr = 0.0
    for s in range(S):
        rr, js = O.xcorr_residual(mod[s], obs[s])
        Jr += js
        assert rel_l2(r[s].double().cpu().numpy(), rr) < 1e-6, s
    assert abs(float(J) - Jr) <= 1e-9 * Jr


def test_exact_gradient_of_the_xcorr_misfit(gpu_lib):
    """Propagator.gradient(misfit='xcorr') against the oracle's exact adjoint
    driven by the oracle's cross-correlation residual."""
    from tests.gpu_helpers import propagator_of
    p = tiny(n=(24, 22, 20), nt=90, W=4, seed=17, nrec=8)
    # data from another mechanism in a perturbed medium: correlations well below 1, so the
    # residual is not a small difference of its two terms (which would amplify the
    # float32 rounding of the modelled traces)
    obs = run_forward(p, m6=np.roll(p.m6, 2), lam=p.lam * 1.1, mu=p.mu * 0.9)[0].astype(np.float32)
    hist = []
    g_ref, _ = run_forward(p, history=hist)
    r, Jr = O.xcorr_residual(g_ref, obs)
    ref = O.exact_adjoint(p.medium(), p.cpml(), p.c, p.src, p.wavelet, p.rec, p.nt, r, hist)
    prop = propagator_of(p)
    J, g, glam, gmu, grho, gm6 = prop.gradient(torch.as_tensor(obs[None], device="cuda"), misfit="xcorr")
    assert abs(J - Jr) <= 1e-4 * abs(Jr)
    for name, got in (("lam", glam), ("mu", gmu), ("rho", grho)):
        e = rel_l2(got.double().cpu().numpy(), ref[name])
        assert e < 1e-4, (name, e)
    assert rel_l2(gm6[0].cpu().numpy(), ref["m6"]) < 1e-4


def test_driver_xcorr_methods(gpu_lib):
    e3, mdl, wav, cpml, geo = _setup(mode=1, nt=100)
    e3.setIdxEarthquake(0)
    e3.setStressTensor(D.StressTensor(1.0, 0.0, 1.0, 0.0, 0.0, 0.0))
    e3.forwardFields("obs")
    e3.setStressTensor(D.StressTensor(0.5, 0.2, 0.5, 0.0, 0.5, 0.0))
    e3.forwardFields("mod")
    e3.computeResidualCrossCorrelation()
    mod, obs = e3.modGather.stacked(), e3.obsGather.stacked()
    rr, _ = O.xcorr_residual(mod, obs)
    assert rel_l2(e3.resGather.stacked(), rr) < 1e-6
    e3.computeCostFunctionCrossCorrelation(0)
    assert e3.costFunction[-1] == pytest.approx(O.xcorr_cost(mod, obs), rel=1e-9)


def test_initial_model_matches_the_reference_recipe(gpu_lib):
    rng = np.random.default_rng(4)
    n = (30, 28, 26)
    tgt = D.Model(n[2], n[1], n[0], 1e3, 1e3, 1e3)
    tgt.Vp = (3000 + 500 * rng.standard_normal((n[1], n[2], n[0]))).astype(np.float32)
    tgt.Vs = (tgt.Vp / 1.73).astype(np.float32)
    tgt.compute_Rho()
    tgt.Rho = np.asarray(tgt.Rho, np.float32)
    tgt.compute_Lambda_Mu()
    cn = D.initial_model(tgt, device="cuda")
    vp = gaussian_filter(tgt.Vp.copy(), sigma=10)
    vp[:, :10, :] = tgt.Vp[:, :10, :]
    np.testing.assert_allclose(cn.Vp, vp, rtol=1e-6)
    from elastic_fm_vp_vs_rho_amd import model as Mo
    rho = Mo.rho_nafe_drake(vp)
    rho[:, :10, :] = tgt.Rho[:, :10, :]
    np.testing.assert_allclose(cn.Rho, rho, rtol=4e-6)   # float32 polynomial, see test_gpu_conditioning
    lam, mu = Mo.lame_from_velocity(cn.Vp, cn.Vs, cn.Rho)
    np.testing.assert_array_equal(cn.Lambda, lam)


def test_frequency_continuation(gpu_lib, tmp_path):
    """Two frequencies of the mode-1 workflow: observed data at each frequency
    saved under the reference's name, each stage's exact-gradient FWI lowers its
    misfit, and the model found at the first frequency starts the second."""
    e3, mdl, wav, cpml, geo = _setup(nq=2, mode=1, n=(34, 32, 36), nt=120)
    tgt = D.Model(mdl.nz, mdl.ny, mdl.nx, mdl.dz, mdl.dy, mdl.dx)
    tgt.Vp, tgt.Vs = (mdl.Vp * np.float32(1.03)), (mdl.Vs * np.float32(1.03))
    tgt.compute_Rho()
    tgt.Rho = np.asarray(tgt.Rho, np.float32)
    tgt.compute_Lambda_Mu()
    cn = D.Model(mdl.nz, mdl.ny, mdl.nx, mdl.dz, mdl.dy, mdl.dx)
    for k in ("Vp", "Vs", "Rho", "Lambda", "Mu"):
        setattr(cn, k, getattr(mdl, k).copy())
    M = D.StressTensor(1.0, 0.0, 1.0, 0.0, 0.3, 0.0)
    f0 = wav.fq
    lam_before = cn.Lambda.copy()
    hist = D.fwi_multifrequency(tgt, cn, geo, [M, M], M, [f0, 1.5 * f0], wav.tprop, 4, 2, dt=wav.dt,
                                device="cuda", results_dir=str(tmp_path), gradient="exact", ite_cte=2,
                                lbfgsLayers=4)
    assert list(hist) == [f0, 1.5 * f0]
    for fq, cost in hist.items():
        assert len(cost) == 2 and np.all(np.isfinite(cost)) and cost[1] < cost[0], (fq, cost)
        arr = np.load(os.path.join(str(tmp_path), "Gathers", f"obsData_{fq}Hz.npy"), allow_pickle=True)
        assert arr.shape == (2,) and arr[0].Vx.shape == (wav.nSamples, geo.sismos[0].getLenStations())
    assert not np.array_equal(cn.Lambda, lam_before)
