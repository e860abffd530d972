"""Host-side driver contracts (CPU): geometry and gather layouts pinned by the
reference's recorded results, the MT step rule, host maps vs the oracle."""
import os

import numpy as np
import pytest

from elastic_fm_vp_vs_rho_amd import driver as D
from elastic_fm_vp_vs_rho_amd import model as Mo
from oracle import elastic_fp64 as O

FIX = os.path.join(os.path.dirname(__file__), "golden", "reference_results", "results_actual.npz")


@pytest.fixture(scope="module")
def ref():
    return np.load(FIX)


def test_station_layout_matches_reference_run(ref):
    """M:2851-2861: x = linspace(20, 79, 20) inner, y = linspace(20, 79, 10)
    outer, z = 2, for each of the 24 earthquakes."""
    st = D.station_grid(100, 100, 20, 20, 10, z=2)
    x = np.array([s[0] for s in st]); y = np.array([s[1] for s in st]); z = np.array([s[2] for s in st])
    n = len(st)
    assert ref["sta_pos_x"].shape == (24 * n,)
    for q in range(24):
        assert np.array_equal(ref["sta_pos_x"][q * n:(q + 1) * n], x)
        assert np.array_equal(ref["sta_pos_y"][q * n:(q + 1) * n], y)
        assert np.array_equal(ref["sta_pos_z"][q * n:(q + 1) * n], z)
    assert np.all(ref["sismos_pos_z"] == 2) and np.all(ref["sismos_pos_y"][-8:] == 52)


def test_geometry_save_reproduces_reference_files(tmp_path, ref):
    g = D.Geometry()
    sx, sy, sz = ref["sismos_pos_x"], ref["sismos_pos_y"], ref["sismos_pos_z"]
    st = D.station_grid(100, 100, 20, 20, 10, z=2)
    for i in range(len(sx)):
        e = D.Earthquake(sx[i], sy[i], sz[i])
        for s in st:
            e.addStation(D.Station(*s))
        g.addSismo(e)
    g.save(str(tmp_path))
    for c in "xyz":
        assert np.array_equal(np.load(tmp_path / f"sta_pos_{c}.npy"), ref[f"sta_pos_{c}"])
        assert np.array_equal(np.load(tmp_path / f"sismos_pos_{c}.npy"), ref[f"sismos_pos_{c}"])


def test_gather_layout_and_nt_rule(ref):
    """Gathers are (nt, nRec) float32, nt = int(tprop / dt) with the hard-coded dt."""
    w = D.Wavelet(0.1, 120.0, None)
    assert tuple(ref["gather_shape"]) == (w.nSamples, 200) == (1403, 200)
    assert bool(ref["gather_dtype_is_f32"])
    gth = D.Gather(w.nSamples, 200)
    assert gth.Vx.shape == (1403, 200) and gth.Vx.dtype == np.float32


def test_mt_first_step_rule(ref):
    """computeAlpha + updateStressTensor: the first MT step is exactly ±0.05 on
    the diagonal and ±0.025 off it (0 where the image is 0), as in every
    recorded trajectory."""
    for q in range(6):
        d = np.array([ref[f"M{k}{q}"][1] - ref[f"M{k}{q}"][0] for k in ("xx", "yy", "zz", "xy", "xz", "yz")])
        for i, v in enumerate(d):
            assert np.isclose(abs(v), 0.05 if i < 3 else 0.025, atol=1e-12) or v == 0.0, (q, i, v)
    e = D.Elastic3D.__new__(D.Elastic3D)
    e.gk = D.Gradients(0, 1, 1, 1)
    e.alpha = D.Alpha()
    e._shots = None
    e.M = D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0)
    rng = np.random.default_rng(0)
    g = rng.standard_normal(6) * 1e-12
    g[4] = 0.0
    e.gk.Mxx, e.gk.Myy, e.gk.Mzz, e.gk.Mxy, e.gk.Mxz, e.gk.Myz = g
    m0 = e.M.as_array()
    e.computeAlpha()
    e.updateStressTensor()
    step = e.M.as_array() - m0
    expect = np.where(g == 0, 0.0, np.sign(g) * np.array([0.05, 0.05, 0.05, 0.025, 0.025, 0.025]))
    np.testing.assert_allclose(step, expect, atol=1e-15)


def test_recorded_costs_never_increase(ref):
    """Trial MTs are kept only when the misfit drops, so each recorded cost history
    is non-increasing (M:2295-2312)."""
    for q in range(6):
        c = ref[f"cost{q}"]
        assert np.all(np.diff(c) <= 0), q
        assert c[-1] < 0.2 * c[0]


def test_host_maps_match_oracle():
    rng = np.random.default_rng(3)
    vp = (2500 + 1500 * rng.random((5, 6, 7))).astype(np.float32)
    vs = (vp / 1.7).astype(np.float32)
    rho = (2000 + 500 * rng.random((5, 6, 7))).astype(np.float32)
    l1, m1 = Mo.lame_from_velocity(vp, vs, rho)
    l2, m2 = O.lame_from_velocity(vp, vs, rho)
    assert np.array_equal(l1, l2) and np.array_equal(m1, m2)
    np.testing.assert_allclose(Mo.rho_nafe_drake(vp), O.rho_nafe_drake(vp), rtol=1e-12)
    for free in (True, False):
        a = Mo.cpml_coefficients(20, 0.1, Mo.REF_DT, 1e3, 4500.0, 100, 90, 80, free_top=free)
        b = O.cpml_coefficients(20, 0.1, Mo.REF_DT, 1e3, 4500.0, 100, 90, 80, free_top=free)
        for k in a:
            np.testing.assert_allclose(a[k], b[k], rtol=2e-7, atol=1e-12)
    np.testing.assert_allclose(Mo.ricker(0.1, Mo.REF_DT, 1403), O.ricker(0.1, Mo.REF_DT, 1403), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(Mo.gaussian(0.2, Mo.REF_DT, 500), O.gaussian(0.2, Mo.REF_DT, 500), rtol=1e-6, atol=1e-7)
