"""Autotuner: every compiled tile reproduces the reference tile's output, and the
gradient computed with the tuned tiles equals the one with the default tiles."""
import pytest
import torch

from tests.gpu_helpers import propagator_of
from tests.problems import tiny

pytestmark = pytest.mark.gpu


def test_all_tiles_parity_clean_and_tuned_gradient_unchanged(gpu_lib):
    from elastic_fm_vp_vs_rho_amd import autotune
    p = tiny(n=(70, 24, 20), nt=60, W=4, seed=41, nrec=6)
    prop = propagator_of(p)
    obs = torch.zeros((1, 3, p.nt, len(p.rec)), device="cuda")
    for k in range(4):
        autotune.set_tile(k, autotune.REF_TILE)
    ref = prop.gradient(obs, K=8)
    rep = autotune.tune(prop, reps=3, use_cache=False)
    for name, r in rep.items():
        rejected = [t for t, v in r["ms"].items() if isinstance(v, str)]
        assert not rejected, (name, rejected)
        assert tuple(r["tile"]) == autotune.get_tile(autotune.KERNELS.index(name))
    got = prop.gradient(obs, K=8)
    for a, b in zip(ref[2:], got[2:]):
        assert float((a.double() - b.double()).abs().max()) <= 1e-6 * float(b.double().abs().max())
    for k in range(4):
        autotune.set_tile(k, autotune.DEFAULT_TILE)
