"""Persistent plans (efm_plan_*): gradients whose forward, segment re-forward
and adjoint loops replay captured HIP graphs give the eager gradients bit for
bit, for every tape strategy, on the first call (capture) and on later calls
(replay, no new graphs), with two shots of different record length.  (The
misfit J is an atomic FP64 sum, so it agrees to rounding, not bit for bit.)"""
import numpy as np
import pytest
import torch

from tests.gpu_helpers import grid_of
from tests.problems import tiny

pytestmark = pytest.mark.gpu


def _two_shot_prop(p, graphs):
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    grid = grid_of(p)
    w = np.stack([p.wavelet, 0.7 * p.wavelet])
    sh = ShotBatch(grid, [p.src, (p.src[0] - 2, p.src[1] + 1, 3)], [p.m6, p.m6[::-1].copy()], w,
                   [p.nt, p.nt - 23], [p.rec, p.rec], p.cp)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    return Propagator(grid, f(p.lam), f(p.mu), f(p.rho), sh, graphs=graphs)


@pytest.mark.parametrize("strategy", ["store", "checkpoint", "hybrid"])
def test_plan_gradient_is_bit_identical_to_eager(gpu_lib, strategy):
    p = tiny(n=(26, 22, 20), nt=97, W=4, seed=51, nrec=6, free_top=True)
    eager = _two_shot_prop(p, graphs=False)
    graph = _two_shot_prop(p, graphs=True)
    obs = torch.as_tensor(np.random.default_rng(3).standard_normal((2, 3, p.nt, 6)).astype(np.float32) * 1e-9,
                          device="cuda")
    m = 40 if strategy == "hybrid" else None
    ref = eager.gradient(obs, K=10, strategy=strategy, m=m)
    counts = []
    for call in range(3):
        got = graph.gradient(obs, K=10, strategy=strategy, m=m)
        torch.cuda.synchronize()
        counts.append(graph.graph_plan().graphs())
        assert abs(got[0] - ref[0]) <= 1e-12 * abs(ref[0]), call    # J: an atomic FP64 sum
        for a, b in zip(got[1:], ref[1:]):
            assert torch.equal(a, b), (strategy, call)
    assert counts[0] > 0 and counts[1] == counts[2] == counts[0], counts   # replays capture nothing new


def test_plan_forward_matches_eager(gpu_lib):
    p = tiny(n=(24, 20, 18), nt=75, W=3, seed=52, nrec=5)
    eager = _two_shot_prop(p, graphs=False)
    graph = _two_shot_prop(p, graphs=True)
    g1, s1 = eager.forward()
    g2, s2 = graph.forward()
    g3, s3 = graph.forward(state=graph.new_wavefield())       # new buffers: a second graph
    torch.cuda.synchronize()
    assert torch.equal(g1, g2) and torch.equal(g1, g3)
    assert torch.equal(s1.fields, s2.fields)


def test_efm_forward_graph_entry_point_matches_eager(gpu_lib):
    """The one-shot graph entry point (captures, replays and releases a graph per
    call, on a non-default stream) still gives the eager forward bit for bit."""
    import ctypes
    from elastic_fm_vp_vs_rho_amd import _capi
    from elastic_fm_vp_vs_rho_amd.propagator import _ptr, _stream
    p = tiny(n=(22, 20, 18), nt=70, W=3, seed=53, nrec=4)
    prop = _two_shot_prop(p, graphs=False)
    g_ref, s_ref = prop.forward()
    st, g = prop.new_wavefield(), prop.new_gathers()
    counter = torch.zeros(1, dtype=torch.int32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    d, m, sh = prop.dims(), prop.medium(), prop.shots.struct()
    with torch.cuda.stream(side):
        _capi.check(gpu_lib.efm_forward_graph(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh),
                                              ctypes.byref(st.struct()), _ptr(g), 0, p.nt, 16, _ptr(counter),
                                              _stream()), "efm_forward_graph")
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert torch.equal(g, g_ref) and torch.equal(st.fields, s_ref.fields)
