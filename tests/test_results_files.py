"""On-disk results in the reference's names, shapes and dtypes (CPU; the
propagation and device kernels replaced by the stand-ins of
test_driver_contract).  The reference writes these under Results/Actual/
(M:2245-2247, 2323-2335, 2414-2417, 2523, 2711-2719); the recorded run's own
files (tests/golden/reference_results) fix the dtypes and the relation between
the lengths of the MT and cost histories."""
import os

import numpy as np
import pytest

from elastic_fm_vp_vs_rho_amd import driver as D
from tests.test_driver_contract import LinearStandIn, _e3d, _fwi_stand_in, _observe

FIX = os.path.join(os.path.dirname(__file__), "golden", "reference_results", "results_actual.npz")
KEYS = ("xx", "yy", "zz", "xy", "xz", "yz")


@pytest.fixture(scope="module")
def ref():
    return np.load(FIX)


def test_fwi_mij_files(tmp_path, ref):
    e3 = _e3d()
    e3.results_dir = str(tmp_path)
    LinearStandIn(e3)
    _observe(e3, D.StressTensor(1.0, 0.0, 1.0, 0.0, 0.0, 0.0))
    e3.setStressTensor(D.StressTensor(0.5, 0.0, 0.5, 0.0, 0.5, 0.0))
    n = 4
    e3.fwi_Mij(n)
    cost = np.load(tmp_path / "CostFunction" / "costFunction0.npy")
    assert cost.dtype == ref["cost0"].dtype == np.float64 and cost.shape == (n,)
    for k in KEYS:
        m = np.load(tmp_path / "TM" / f"Mmod{k}0.npy")
        # the recorded run: MT history one longer than the cost history
        assert m.dtype == ref[f"M{k}0"].dtype and m.shape == (n + 1,)
        assert ref[f"M{k}0"].shape[0] == ref["cost0"].shape[0] + 1
        assert m[0] == {"xx": 0.5, "yy": 0.0, "zz": 0.5, "xy": 0.0, "xz": 0.5, "yz": 0.0}[k]
    for stem in ("modData0", "resData0", "modDataIni0"):
        for c in ("Vx", "Vy", "Vz"):
            a = np.load(tmp_path / "Gathers" / f"{stem}.{c}.npy")
            assert a.dtype == np.float32 and a.shape == (e3.nt, e3.LenStations), (stem, c)
    assert bool(ref["gather_dtype_is_f32"]) and len(ref["gather_shape"]) == 2


def test_fwi_vp_vs_rho_files(tmp_path):
    tm = [D.StressTensor(1, 0, 1, 0, 0, 0)]
    e3 = _e3d(mode=1)
    e3.results_dir = str(tmp_path)
    _fwi_stand_in(e3, 7, value=9.0)                  # every trial improves
    n = 3
    cost = e3.fwi_Vp_Vs_Rho([None], tm, n, lbfgsLayers=4, ite_cte=1)
    fq = e3.ondicula.fq
    shape = (e3.ny, e3.nz, e3.nx)
    for name in ("Vp", "Vs", "Rho", "Lambda", "Mu"):
        a = np.load(tmp_path / "Models" / f"{name}_{fq}Hz.npy")
        assert a.dtype == np.float32 and a.shape == shape, name
        np.testing.assert_array_equal(a, getattr(e3.mk, name))
    c = np.load(tmp_path / "CostFunction" / f"costFunction_{fq}Hz.npy")
    assert c.dtype == np.float64 and c.shape == (n,)
    np.testing.assert_array_equal(c, cost)
    for name in ("gRho", "gMu", "gLambda"):
        g = np.load(tmp_path / "Gradients" / f"{name}_{fq}Hz.npy")
        assert g.dtype == np.float32 and g.shape == shape, name
    # the list of per-earthquake Gather objects (a pickled object array, as np.save
    # writes the reference's list; our own file, read back with allow_pickle)
    mod = np.load(tmp_path / "Gathers" / f"modData_{fq}Hz.npy", allow_pickle=True)
    assert mod.dtype == object and mod.shape == (e3.geometria.getLenSismos(),)
    for d in mod:
        for comp in ("Vx", "Vy", "Vz"):
            v = getattr(d, comp)
            assert v.dtype == np.float32 and v.shape == (e3.nt, e3.LenStations)
