"""Drop-in driver on the GPU against the FP64 oracle: the reference-style
forward, MT adjoint (backwardFields) and λ/μ/ρ imaging with reverse-time
reconstruction (backwardFields_FowardFieldsRemake), and the inversion drivers'
contracts (first-step rule, monotone accepted misfit, NaN path)."""
import numpy as np
import pytest

from elastic_fm_vp_vs_rho_amd import driver as D
from oracle import elastic_fp64 as O
from tests.problems import rel_l2

pytestmark = pytest.mark.gpu


def _setup(nq=1, mode=0, n=(26, 24, 22), W=4, nt=140, seed=5):
    nx, ny, nz = n
    rng = np.random.default_rng(seed)
    mdl = D.Model(nz, ny, nx, 10.0, 10.0, 10.0)
    mdl.Vp = (3000 + 300 * rng.standard_normal((ny, nz, nx))).astype(np.float32)
    mdl.Vs = (mdl.Vp / 1.75).astype(np.float32)
    mdl.Rho = (2300 + 100 * rng.random((ny, nz, nx))).astype(np.float32)
    mdl.compute_Lambda_Mu()
    mdl.MaxMinVp()
    dt = 0.4 * 10.0 / float(mdl.VmaxVp)
    wav = D.Wavelet(0.15 / dt / 8, (nt + 0.5) * dt, mdl, dt=dt)
    wav.ricker()
    cpml = D.CPML(W, wav, mdl, free_top=True)
    geo = D.Geometry()
    stations = D.station_grid(nx, ny, W, 4, 3, z=2)
    for q in range(nq):
        e = D.Earthquake(int(rng.integers(W + 2, nx - W - 2)), int(rng.integers(W + 2, ny - W - 2)), 2)
        for s in stations:
            e.addStation(D.Station(*s))
        geo.addSismo(e)
    e3 = D.Elastic3D(mdl, geo, wav, cpml, mode=mode)
    return e3, mdl, wav, cpml, geo


def _oracle(mdl, wav, cpml, geo, q=0):
    med = O.Medium(mdl.Lambda, mdl.Mu, mdl.Rho)
    cp = O.Cpml.from_dict(cpml.profiles())
    eq = geo.sismos[q]
    rec = np.array([(s.x, s.y, s.z) for s in eq.stations])
    return med, cp, (eq.x, eq.y, eq.z), rec


def test_reference_forward_and_mt_adjoint(gpu_lib):
    e3, mdl, wav, cpml, geo = _setup(mode=0)
    M_obs = D.StressTensor(1.0, 0.0, 1.0, 0.0, 0.0, 0.0)
    M_mod = D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0)
    e3.setIdxEarthquake(0)
    e3.setStressTensor(M_obs)
    e3.forwardFields("obs")
    med, cp, src, rec = _oracle(mdl, wav, cpml, geo)
    w = wav.waveform[:, 0]
    c = wav.dt / mdl.dx
    g_obs, _ = O.forward(med, cp, c, src, M_obs.as_array(), w, rec, wav.nSamples)
    got = np.stack(e3.getGathers())
    assert rel_l2(got, g_obs) < 1e-5
    np.testing.assert_array_equal(e3.obsGather.Vx, got[0])
    e3.setStressTensor(M_mod)
    e3.forwardFields("mod")
    e3.computeResidual()
    e3.backwardFields()
    g_mod, _ = O.forward(med, cp, c, src, M_mod.as_array(), w, rec, wav.nSamples)
    res = g_mod - g_obs
    ref = O.ref_backward_mt(med, cp, c, src, w, rec, wav.nSamples, res)
    gk = np.array([e3.gk.Mxx, e3.gk.Myy, e3.gk.Mzz, e3.gk.Mxy, e3.gk.Mxz, e3.gk.Myz])
    assert rel_l2(gk, ref) < 1e-4, (gk, ref)


def test_reference_remake_imaging(gpu_lib):
    e3, mdl, wav, cpml, geo = _setup(mode=1, seed=6)
    M = D.StressTensor(0.8, -0.3, 0.6, 0.2, 0.1, -0.4)
    e3.setIdxEarthquake(0)
    e3.setStressTensor(M)
    med, cp, src, rec = _oracle(mdl, wav, cpml, geo)
    w = wav.waveform[:, 0]
    c = wav.dt / mdl.dx
    nt = wav.nSamples
    obs = O.forward(O.Medium(mdl.Lambda * 1.05, mdl.Mu, mdl.Rho), cp, c, src, M.as_array(), w, rec, nt)[0]
    e3.obsGather.Vx, e3.obsGather.Vy, e3.obsGather.Vz = (obs[k].astype(np.float32) for k in range(3))
    e3.forwardFields("mod")
    e3.computeResidual()
    e3.backwardFields_FowardFieldsRemake()
    borders = {}
    g_mod, Ff = O.forward(med, cp, c, src, M.as_array(), w, rec, nt, borders=borders, W=cpml.lenPML)
    res = g_mod - obs.astype(np.float32)
    glam, gmu, grho = O.ref_backward_remake(med, cp, c, src, M.as_array(), w, rec, nt, res, Ff, borders,
                                            cpml.lenPML, wav.dt, mdl.dx)
    for name, got, want in (("lam", e3.gk.Lambda, glam), ("mu", e3.gk.Mu, gmu), ("rho", e3.gk.Rho, grho)):
        e = rel_l2(got, want)
        assert e < 1e-4, (name, e)


def test_host_driven_loops_match_the_c_loops(gpu_lib):
    """host_loop: the reference's structure (one host call per step, each
    iteration's three images copied to the host and summed there) gives the C
    loops' gathers bit for bit and their images to float32 summation order."""
    out = {}
    for host_loop in (False, True):
        e3, mdl, wav, cpml, geo = _setup(mode=1, seed=8, nt=90)
        e3.host_loop = host_loop
        e3.setIdxEarthquake(0)
        e3.setStressTensor(D.StressTensor(0.7, 0.2, -0.4, 0.1, 0.3, -0.2))
        rng = np.random.default_rng(2)
        e3.obsGather.Vx, e3.obsGather.Vy, e3.obsGather.Vz = (
            (1e-9 * rng.standard_normal((wav.nSamples, e3.LenStations))).astype(np.float32) for _ in range(3))
        e3.forwardFields("mod")
        e3.computeResidual()
        e3.backwardFields_FowardFieldsRemake()
        out[host_loop] = (np.stack(e3.getGathers()), e3.gk.Lambda.copy(), e3.gk.Mu.copy(), e3.gk.Rho.copy())
    np.testing.assert_array_equal(out[True][0], out[False][0])
    for k in (1, 2, 3):
        assert np.abs(out[False][k]).max() > 0
        assert rel_l2(out[True][k], out[False][k]) < 1e-6, k


def test_fwi_mij_contract(gpu_lib):
    e3, mdl, wav, cpml, geo = _setup(mode=0, nt=100, seed=7)
    e3.setIdxEarthquake(0)
    e3.setStressTensor(D.StressTensor(1.0, 0.0, 1.0, 0.0, 0.0, 0.0))
    e3.forwardFields("obs")
    e3.setStressTensor(D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0))
    hist = e3.fwi_Mij(5)
    steps = np.array([hist[k][1] - hist[k][0] for k in ("xx", "yy", "zz", "xy", "xz", "yz")])
    for i, v in enumerate(steps):
        assert np.isclose(abs(v), 0.05 if i < 3 else 0.025) or v == 0.0
    assert np.all(np.diff(e3.costFunction) <= 0)
    assert e3.costFunction[-1] < e3.costFunction[0]


def test_fwi_mij_nan_path_reverts_and_halves(gpu_lib):
    e3, mdl, wav, cpml, geo = _setup(mode=0, nt=80, seed=8)
    e3.setIdxEarthquake(0)
    e3.setStressTensor(D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0))
    e3.forwardFields("obs")
    e3.obsGather.Vx[3, 1] = np.nan                          # forced fault: NaN observed sample
    hist = e3.fwi_Mij(2)
    a = e3.alpha
    assert np.isnan(e3.costFunction[1])
    # iteration 1 rejected its trial: M kept, every α halved
    assert all(hist[k][2] == hist[k][1] for k in hist)
    e3.computeAlpha()   # recompute from the (NaN-free) last image for comparison
    assert np.isfinite(a.Mxx)


@pytest.mark.parametrize("gradient", ["reference", "exact"])
def test_fwi_vp_vs_rho_runs_and_descends(gpu_lib, gradient):
    e3, mdl, wav, cpml, geo = _setup(nq=2, mode=1, n=(30, 28, 26), W=4, nt=120, seed=9)
    # the model update re-derives ρ from Vp (Nafe-Drake), so start from such a model
    # as the reference run does (cn.compute_Rho(), M:3072-3073)
    mdl.compute_Rho()
    mdl.Rho = np.asarray(mdl.Rho, np.float32)
    mdl.compute_Lambda_Mu()
    e3.setModel(mdl)
    tm = [D.StressTensor(1.0, 0.2, 1.0, 0.0, 0.3, 0.0)] * 2
    true = D.Model(mdl.nz, mdl.ny, mdl.nx, mdl.dz, mdl.dy, mdl.dx)
    true.Vp, true.Vs = mdl.Vp * 1.03, mdl.Vs * 1.03
    true.compute_Rho()
    true.Rho = np.asarray(true.Rho, np.float32)
    true.compute_Lambda_Mu()
    e3.setModel(true)
    obs = []
    for q in range(2):
        e3.setIdxEarthquake(q)
        e3.setStressTensor(tm[q])
        e3.forwardFields("obs")
        g = D.Gather(wav.nSamples, e3.LenStations)
        g.Vx, g.Vy, g.Vz = e3.getGathers()
        obs.append(g)
    e3.setModel(mdl)
    cost = e3.fwi_Vp_Vs_Rho(obs, tm, 2, lbfgsLayers=4, ite_cte=2, gradient=gradient)
    assert len(cost) == 2 and all(np.isfinite(cost))
    if gradient == "exact":
        assert cost[1] < cost[0]


def test_exact_gradient_through_the_driver_reuses_its_states(gpu_lib):
    """Elastic3D.gradient_exact matches the FP64 oracle's exact gradient; a second
    call (new model, same earthquake) reuses the kept forward states and still
    matches; release_states frees them."""
    e3, mdl, wav, cpml, geo = _setup(mode=1, nt=90)
    e3.setIdxEarthquake(0)
    e3.setStressTensor(D.StressTensor(1.0, 0.0, 1.0, 0.0, 0.0, 0.0))
    e3.forwardFields("obs")
    e3.setStressTensor(D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0))
    med, cp, src, rec = _oracle(mdl, wav, cpml, geo)
    w = wav.waveform[:, 0]
    c = wav.dt / mdl.dx
    obs = np.stack([e3.obsGather.Vx, e3.obsGather.Vy, e3.obsGather.Vz]).astype(np.float64)
    for k in range(2):
        if k == 1:                                   # a new model: the states are reused
            mdl2 = D.Model(mdl.nz, mdl.ny, mdl.nx, mdl.dz, mdl.dy, mdl.dx)
            mdl2.Vp, mdl2.Vs, mdl2.Rho = mdl.Vp * np.float32(1.01), mdl.Vs.copy(), mdl.Rho.copy()
            mdl2.compute_Lambda_Mu()
            e3.setModel(mdl2)
            med = O.Medium(mdl2.Lambda, mdl2.Mu, mdl2.Rho)
        J, g = e3.gradient_exact()
        assert len(e3._state_pool) == wav.nSamples + 1
        hist = []
        gr, _ = O.forward(med, cp, c, src, e3.M.as_array(), w, rec, wav.nSamples, history=hist)
        r, Jr = O.misfit(gr, obs)
        ref = O.exact_adjoint(med, cp, c, src, w, rec, wav.nSamples, r, hist)
        assert abs(J - Jr) <= 1e-4 * Jr
        for name in ("lam", "mu", "rho"):
            e = rel_l2(g[name].double().cpu().numpy(), ref[name])
            assert e < 1e-4, (k, name, e)
        assert rel_l2(g["m6"], ref["m6"]) < 1e-4
    e3.release_states()
    assert not e3._state_pool
