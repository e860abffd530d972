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

from elastic_fm_vp_vs_rho_amd import driver as D

KEYS = ("xx", "yy", "zz", "xy", "xz", "yz")


def _e3d(mode=0, nt=40):
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
    e3 = D.Elastic3D(mdl, geo, wav, cpml, mode=mode, device="cpu")
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


def _fwi_stand_in(e3, seed):
    rng = np.random.default_rng(seed)
    calls = {"misfit": 0}

    def all_shots(data_obs, listaTM, gradient):
        g = lambda: rng.standard_normal((e3.ny, e3.nz, e3.nx))
        return 10.0, g(), g(), g(), [None]

    def misfit_all(data_obs, listaTM):
        calls["misfit"] += 1
        return 11.0                                               # never better than 10

    e3._all_shots_gradient = all_shots
    e3._misfit_all = misfit_all
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
