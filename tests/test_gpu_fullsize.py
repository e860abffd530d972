"""Full-size workloads on the GPU.

C1 (64³, 1000 steps): seismograms against the committed FP64 oracle fixture
(tests/golden/make_c1_fixture.py), on the eager and the hipGraph paths, and
linearity in the moment tensor.  C3 (256×256×128, 32 shots, uneven record
lengths): every sampled shot of the batched run equals its single-shot run bit
for bit, and nothing is recorded past a shot's own record length."""
import os

import numpy as np
import pytest
import torch

from tests.problems import rel_l2

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "c1_gathers.npz")


def _prop(wl, m6=None, shots_idx=None):
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, Propagator, ShotBatch, lame
    dev = torch.device("cuda")
    grid = Grid(wl.nx, wl.ny, wl.nz, wl.dx, wl.dt, (wl.W, wl.W, wl.W, wl.W, 0 if wl.free_top else wl.W, wl.W))
    idx = np.arange(wl.S) if shots_idx is None else np.asarray(shots_idx)
    cp = {k: v[idx] for k, v in wl.cpml().items()}
    wav = wl.wavelets()[idx]
    shots = ShotBatch(grid, wl.src[idx], wl.m6[idx] if m6 is None else m6, wav, wl.nt[idx], wl.rec[idx], cp,
                      device=dev)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    lam, mu = lame(f(wl.vp), f(wl.vs), f(wl.rho))
    return Propagator(grid, lam, mu, f(wl.rho), shots)


def test_c1_against_fp64_fixture(gpu_lib):
    from elastic_fm_vp_vs_rho_amd import workloads
    wl = workloads.c1()
    fx = np.load(GOLD)
    assert np.allclose(fx["m6"], wl.m6[0]) and np.array_equal(fx["rec"], wl.rec[0])
    prop = _prop(wl)
    g, _ = prop.forward()
    e = rel_l2(g[0].double().cpu().numpy(), fx["gathers"])
    assert e < 1e-5, e
    gg, _ = prop.forward(graph_chunk=50)
    torch.cuda.synchronize()
    assert torch.equal(gg, g)


def test_c1_linearity_in_moment_tensor(gpu_lib):
    from elastic_fm_vp_vs_rho_amd import workloads
    wl = workloads.c1()
    rng = np.random.default_rng(5)
    m1, m2 = rng.uniform(-1, 1, 6), rng.uniform(-1, 1, 6)
    g1 = _prop(wl, m6=m1[None]).forward()[0].double()
    g2 = _prop(wl, m6=m2[None]).forward()[0].double()
    g12 = _prop(wl, m6=(m1 + m2)[None]).forward()[0].double()
    # three independent FP32 runs of 1000 steps: linear to FP32 accumulation
    # accuracy (each run alone is within 1e-5 of FP64, see the fixture test)
    assert float(torch.linalg.norm(g1 + g2 - g12) / torch.linalg.norm(g12)) < 2e-5


def test_c3_batched_equals_single_shots(gpu_lib):
    from elastic_fm_vp_vs_rho_amd import workloads
    wl = workloads.c3()
    assert wl.S == 32 and wl.nt.min() >= 1500 and wl.nt.max() <= 3000
    batch = _prop(wl)
    gb, _ = batch.forward(graph_chunk=100)
    torch.cuda.synchronize()
    assert torch.isfinite(gb).all()
    order = np.argsort(wl.nt)
    for s in (int(order[0]), int(order[-1]), 7):
        g1, _ = _prop(wl, shots_idx=[s]).forward()
        n = int(wl.nt[s])
        assert torch.equal(gb[s, :, :n], g1[0, :, :n]), s
        assert not torch.any(gb[s, :, n:]), s
        assert float(g1.abs().max()) > 0
