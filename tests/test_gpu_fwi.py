"""Batched exact-gradient FWI + MT (workload C4 at a reduced size): the gradient
summed over sub-batches equals the one-shot-at-a-time gradient, and the joint
L-BFGS inversion of Vp, Vs, ρ and the moment tensors lowers the misfit at every
accepted iteration."""
import numpy as np
import pytest
import torch

from tests.problems import rel_l2

pytestmark = pytest.mark.gpu


def _small():
    from elastic_fm_vp_vs_rho_amd import fwi, workloads
    kw = dict(nx=48, ny=44, nz=40, S=5, nt=160, nrec=16)
    wt, w0 = workloads.c4(anomalies=True, **kw), workloads.c4(anomalies=False, **kw)
    dev = torch.device("cuda")
    survey = fwi.Survey(w0, dev)
    f = lambda a: fwi._tensor(a, dev)
    obs = fwi.observed(survey, f(wt.vp), f(wt.vs), f(wt.rho), wt.m6, 2)
    return fwi, survey, f, wt, w0, obs


def test_batched_gradient_equals_one_shot_at_a_time(gpu_lib):
    fwi, survey, f, wt, w0, obs = _small()
    m6 = np.tile([0.5, 0.0, 0.5, 0.0, 0.5, 0.0], (w0.S, 1))
    args = (survey, f(w0.vp), f(w0.vs), f(w0.rho), m6, obs)
    Jp, *gp = fwi.gradient(*args, 1)
    Jb, *gb = fwi.gradient(*args, 3)
    assert abs(Jb - Jp) <= 1e-12 * Jp
    for a, b in zip(gb[:3], gp[:3]):
        assert rel_l2(a.double().cpu().numpy(), b.double().cpu().numpy()) < 1e-6
    np.testing.assert_array_equal(gb[3], gp[3])          # per-shot MT gradients do not mix shots
    assert survey.max_batch() >= 1


def test_joint_inversion_lowers_the_misfit(gpu_lib):
    fwi, survey, f, wt, w0, obs = _small()
    m6 = np.tile([0.5, 0.0, 0.5, 0.0, 0.5, 0.0], (w0.S, 1))
    hist, (vp, vs, rho, m6_end) = fwi.invert(survey, f(w0.vp), f(w0.vs), f(w0.rho), m6, obs, 3, 3)
    assert len(hist) == 4 and np.all(np.isfinite(hist))
    assert np.all(np.diff(hist) < 0), hist
    assert hist[-1] < 0.5 * hist[0], hist
    # nothing changes inside the absorbing layers
    mask = fwi.update_mask(survey.grid).to(vp.device)
    assert torch.equal(vp * (1 - mask), f(w0.vp) * (1 - mask))
