"""Driver control flow on the CPU, with the propagation calls of Elastic3D
replaced by a linear stand-in (d = G·M and its transpose as the MT image), so
that the reference's inversion rules are exercised without a GPU:

* fwi_Mij (M:2217-2346): first step ±0.05 on the diagonal and ±0.025 off it,
  a trial is accepted only when the misfit decreases, otherwise every α is
  halved and M is kept; a NaN misfit rejects the trial;
* fwi_Vp_Vs_Rho (M:2547-2664): when no trial of the halving line search lowers
  the misfit, the model is restored.
"""
import numpy as np
import pytest

from elastic_fm_vp_vs_rho_amd import driver as D

KEYS = ("xx", "yy", "zz", "xy", "xz", "yz")


def _e3d(mode=0, nt=40, device="cpu"):
    # large enough that the FWI driver's masks (top 10 cells, CPML + 2) leave cells to update
    nx, ny, nz = (20, 18, 16) if mode == 0 else (30, 28, 32)
    mdl = D.Model(nz, ny, nx, 10.0, 10.0, 10.0)
    mdl.Vp = np.full((ny, nz, nx), 3000.0, np.float32)
    mdl.Vs = (mdl.Vp / 1.75).astype(np.float32)
    mdl.Rho = np.full((ny, nz, nx), 2400.0, np.float32)
    mdl.compute_Lambda_Mu()
    mdl.MaxMinVp()
    dt = 0.4 * 10.0 / 3000.0
    wav = D.Wavelet(10.0, (nt + 0.5) * dt, mdl, dt=dt)
    wav.ricker()
    cpml = D.CPML(4, wav, mdl)
    geo = D.Geometry()
    eq = D.Earthquake(10, 9, 6)
    for s in D.station_grid(nx, ny, 4, 3, 2, z=2):
        eq.addStation(D.Station(*s))
    geo.addSismo(eq)
    e3 = D.Elastic3D(mdl, geo, wav, cpml, mode=mode, device=device)
    e3.setIdxEarthquake(0)
    return e3


class LinearStandIn:
    """Seismograms linear in M: d = Σ_k M_k G_k; the MT image is the descent
    direction -Gᵀ r of J = ½‖d - obs‖² (the reference adds α·image)."""

    def __init__(self, e3, seed=0):
        self.e3 = e3
        rng = np.random.default_rng(seed)
        self.G = rng.standard_normal((6, 3, e3.nt, e3.LenStations)).astype(np.float32)
        e3.forwardFields = self.forward
        e3.backwardFields = self.backward

    def forward(self, strname="mod", saveFieldFlag=False):
        e3 = self.e3
        d = np.tensordot(e3.M.as_array(), self.G, 1).astype(np.float32)
        e3.gather.Vx, e3.gather.Vy, e3.gather.Vz = d[0], d[1], d[2]
        tgt = e3.modGather if strname == "mod" else e3.obsGather
        tgt.Vx, tgt.Vy, tgt.Vz = e3.getGathers()

    def backward(self, saveFieldFlag=False):
        e3 = self.e3
        r = e3.resGather.stacked().astype(np.float64)
        img = -np.tensordot(self.G.astype(np.float64), r, axes=([1, 2, 3], [0, 1, 2]))
        e3.gk.Mxx, e3.gk.Myy, e3.gk.Mzz, e3.gk.Mxy, e3.gk.Mxz, e3.gk.Myz = (float(v) for v in img)


def _observe(e3, M):
    e3.setStressTensor(M)
    e3.forwardFields("obs")


def test_fwi_mij_first_step_and_monotone_cost():
    e3 = _e3d()
    LinearStandIn(e3)
    _observe(e3, D.StressTensor(1.0, 0.0, 1.0, 0.0, 0.0, 0.0))
    e3.setStressTensor(D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0))      # the reference's start (M:2777)
    hist = e3.fwi_Mij(12)
    first = np.array([hist[k][1] - hist[k][0] for k in KEYS])
    np.testing.assert_allclose(np.abs(first), [0.05, 0.05, 0.05, 0.025, 0.025, 0.025], rtol=1e-6)
    cost = np.array(e3.costFunction)
    assert np.all(np.diff(cost) <= 0) and cost[-1] < cost[0]


def test_fwi_mij_rejected_trial_halves_every_alpha_and_keeps_m():
    e3 = _e3d()
    stand_in = LinearStandIn(e3)
    _observe(e3, D.StressTensor(1.0, 0.2, 1.0, 0.1, 0.0, -0.3))
    e3.setStressTensor(D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0))
    calls = {"n": 0, "alpha": None}

    def overshooting_image(saveFieldFlag=False):
        # from the second iteration on the image is scaled so that the trial overshoots
        stand_in.backward()
        calls["n"] += 1
        if calls["n"] == 2:
            calls["alpha"] = np.array([getattr(e3.alpha, "M" + k) for k in KEYS])
            for k in KEYS:
                setattr(e3.gk, "M" + k, 1e6 * getattr(e3.gk, "M" + k))

    e3.backwardFields = overshooting_image
    hist = e3.fwi_Mij(2)
    assert all(hist[k][2] == hist[k][1] for k in KEYS)              # M kept
    a2 = np.array([getattr(e3.alpha, "M" + k) for k in KEYS])
    np.testing.assert_allclose(a2, calls["alpha"] / 2)               # every α halved
    assert e3.costFunction[1] < e3.costFunction[0]


def test_fwi_mij_nan_cost_rejects_the_trial():
    """A trial whose forward run blows up (NaN misfit) is rejected like a worse one."""
    e3 = _e3d()
    stand_in = LinearStandIn(e3)
    _observe(e3, D.StressTensor(1.0, 0.0, 1.0, 0.0, 0.0, 0.0))
    e3.setStressTensor(D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0))

    def unstable_forward(strname="mod", saveFieldFlag=False):
        stand_in.forward(strname)
        if np.abs(e3.M.as_array()).max() > 5.0:                  # forced fault: the run diverges
            e3.gather.Vx[:] = np.nan
            e3.modGather.Vx = e3.gather.Vx.copy()

    calls = {"n": 0}

    def large_image(saveFieldFlag=False):
        stand_in.backward()
        calls["n"] += 1
        if calls["n"] >= 2:
            for k in KEYS:
                setattr(e3.gk, "M" + k, 1e6 * getattr(e3.gk, "M" + k))

    e3.forwardFields = unstable_forward
    e3.backwardFields = large_image
    hist = e3.fwi_Mij(3)
    assert np.all(np.isfinite(e3.costFunction))
    assert all(hist[k][3] == hist[k][2] == hist[k][1] for k in KEYS)   # both trials rejected
    assert np.all(np.isfinite([hist[k][-1] for k in KEYS]))


def _host_condition(e3, g, top=10):
    """The reference's conditioning (M:2479-2504) with scipy, for host stand-ins."""
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(np.asarray(g, np.float32), sigma=1)
    W = e3.lenPML
    g[:, :top, :] = 0
    g[:W + 2] = 0
    g[e3.ny - W - 2:] = 0
    g[:, :, :W + 2] = 0
    g[:, :, e3.nx - W - 2:] = 0
    g[:, e3.nz - W - 2:, :] = 0
    return g


def _fwi_stand_in(e3, seed, value=11.0):
    """Stand-ins for the propagation and for the device kernels (conditioning,
    model update), so that the driver's control flow runs on the CPU."""
    import torch
    from elastic_fm_vp_vs_rho_amd import model as Mo
    rng = np.random.default_rng(seed)
    calls = {"misfit": 0, "apply": []}

    def all_shots(data_obs, listaTM, gradient):
        g = lambda: torch.as_tensor(rng.standard_normal((e3.ny, e3.nz, e3.nx)).astype(np.float32))
        mod = []
        for _ in range(e3.geometria.getLenSismos()):
            d = D.Gather(e3.nt, e3.LenStations)
            d.Vx[:] = rng.standard_normal(d.Vx.shape)
            mod.append(d)
        return 10.0, g(), g(), g(), mod

    def misfit_all(data_obs, listaTM):
        calls["misfit"] += 1
        return value

    def condition(g, sigma=1.0, top=10):
        return torch.as_tensor(_host_condition(e3, g.numpy(), top))

    def apply_model(lam0, dlam, alam, mu0, dmu, amu, cRho):
        calls["apply"].append((alam, amu, float(dlam.numpy().max()), float(dmu.numpy().max())))
        lam = (lam0.numpy().astype(np.float64) + alam * dlam.numpy()).astype(np.float32)
        mu = (mu0.numpy().astype(np.float64) + amu * dmu.numpy()).astype(np.float32)
        vp, vs = Mo.velocity_from_lame(lam, mu, e3.mk.d_Rho.numpy())
        rho = Mo.rho_nafe_drake(vp).astype(np.float32)
        rho[:, :10, :] = cRho.numpy()[:, :10, :]
        e3._set_model_device(*(torch.as_tensor(np.ascontiguousarray(a, np.float32)) for a in (lam, mu, vp, vs, rho)))

    e3._all_shots_gradient = all_shots
    e3._misfit_all = misfit_all
    e3._condition = condition
    e3._apply_model = apply_model
    return calls


def test_fwi_vp_vs_rho_restores_the_model_when_no_trial_helps():
    tm = [D.StressTensor(1, 0, 1, 0, 0, 0)]
    e3 = _e3d(mode=1)
    calls = _fwi_stand_in(e3, 3)
    cost = e3.fwi_Vp_Vs_Rho([None], tm, 2, lbfgsLayers=4, ite_cte=1, nTry=4)
    assert len(cost) == 2 and calls["misfit"] == 4                # nTry trials in iteration 1
    # every trial of iteration 1 failed: the model is the one iteration 0 produced
    e3b = _e3d(mode=1)
    _fwi_stand_in(e3b, 3)
    e3b.fwi_Vp_Vs_Rho([None], tm, 1, lbfgsLayers=4, ite_cte=1, nTry=4)
    assert np.all(np.isfinite(e3.mk.Lambda)) and np.all(np.isfinite(e3.mk.Mu))
    np.testing.assert_array_equal(e3.mk.Lambda, e3b.mk.Lambda)
    np.testing.assert_array_equal(e3.mk.Mu, e3b.mk.Mu)
    assert not np.array_equal(e3b.mk.Lambda, _e3d(mode=1).mk.Lambda)   # iteration 0 did move it


def test_fwi_vp_vs_rho_steepest_steps_follow_the_reference_alphas():
    """The fixed-α iterations (M:2669-2686) move λ by 1e-5 and μ by 1e-4 of the
    model's 2-norm at the image maximum: α·max(image) = fraction·‖model‖₂."""
    tm = [D.StressTensor(1, 0, 1, 0, 0, 0)]
    e3 = _e3d(mode=1)
    lam0, mu0 = e3.mk.Lambda.astype(np.float64), e3.mk.Mu.astype(np.float64)
    calls = _fwi_stand_in(e3, 5)
    e3.fwi_Vp_Vs_Rho([None], tm, 1, lbfgsLayers=4, ite_cte=2)
    (alam, amu, mlam, mmu), = calls["apply"]
    np.testing.assert_allclose(alam * mlam, 1e-5 * np.linalg.norm(lam0), rtol=1e-6)
    np.testing.assert_allclose(amu * mmu, 1e-4 * np.linalg.norm(mu0), rtol=1e-6)
    assert D.steepest_alphas(lam0, mu0, e3.mk.Rho, np.full(3, 2.0), np.full(3, 4.0), np.full(3, 8.0))[2] \
        == pytest.approx(1e-6 * np.linalg.norm(e3.mk.Rho.astype(np.float64)) / 8.0)
    # the update keeps the top 10 cells of ρ and moves the rest by Nafe-Drake of the new Vp
    from elastic_fm_vp_vs_rho_amd import model as Mo
    np.testing.assert_allclose(e3.mk.Rho[:, 10:, :], Mo.rho_nafe_drake(e3.mk.Vp)[:, 10:, :], rtol=1e-6)


def test_host_condition_stand_in_is_the_reference_rule():
    """The CPU stand-in used above zeroes exactly z < 10 and the CPML+2 zone and
    smooths with scipy's σ=1 Gaussian (mode='reflect', truncate=4)."""
    from scipy.ndimage import gaussian_filter, correlate1d
    e3 = _e3d(mode=1)
    g = np.random.default_rng(0).standard_normal((e3.ny, e3.nz, e3.nx)).astype(np.float32)
    out = _host_condition(e3, g)
    W = e3.lenPML
    keep = np.zeros_like(out, bool)
    keep[W + 2:e3.ny - W - 2, 10:e3.nz - W - 2, W + 2:e3.nx - W - 2] = True
    assert not np.any(out[~keep])
    x = np.arange(-4, 5)
    w = np.exp(-0.5 * x ** 2)
    w /= w.sum()
    ref = g
    for ax in range(3):
        ref = correlate1d(ref, w, axis=ax, mode="reflect")
    np.testing.assert_allclose(out[keep], ref[keep], rtol=1e-6, atol=1e-7)
    np.testing.assert_array_equal(out[keep], gaussian_filter(g, 1.0, mode="reflect", truncate=4.0)[keep])
