"""GPU parity of the HIP propagator against the FP64 oracle (called through the C ABI).

Tolerances (relative L2, FP32 device vs FP64 oracle on identical float32 inputs):
one step 1e-6 per field; seismograms 1e-5; exact gradients 1e-4 per field.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import elastic_fp64 as O
from tests.gpu_helpers import propagator_of, random_state, state_from_device, state_to_device
from tests.problems import rel_l2, run_forward, tiny

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4


def _step_gpu(prop, wf, it):
    from elastic_fm_vp_vs_rho_amd import _capi
    from elastic_fm_vp_vs_rho_amd.propagator import _stream
    L = _capi.lib()
    d, m, sh = prop.dims(), prop.medium(), prop.shots.struct()
    s = wf.struct()
    _capi.check(L.efm_stress_step(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh), ctypes.byref(s),
                                  ctypes.byref(s), it, _stream()), "stress")
    _capi.check(L.efm_velocity_step(ctypes.byref(d), ctypes.byref(m), ctypes.byref(sh), ctypes.byref(s),
                                    ctypes.byref(s), it, _stream()), "velocity")


@pytest.mark.parametrize("free_top", [False, True])
def test_one_step_from_random_state(gpu_lib, free_top):
    p = tiny(n=(70, 13, 12), W=3, seed=3, free_top=free_top)
    rng = np.random.default_rng(8)
    F = random_state(p, rng)
    prop = propagator_of(p)
    wf = prop.new_wavefield()
    state_to_device(F, wf)
    _step_gpu(prop, wf, 5)
    torch.cuda.synchronize()
    G = state_from_device(wf, F.shape)
    med, cp = p.medium(), p.cpml()
    O.stress_sweep(F, med, cp, p.c)
    O.add_source(F, p.src, p.m6, float(p.wavelet[5]))
    O.velocity_sweep(F, med, cp, p.c)
    for k in O.FIELDS:
        assert rel_l2(G.f[k], F.f[k]) < 1e-6, k
    for k in O.PSI_V + O.PSI_T:
        if np.any(F.psi[k]):
            assert rel_l2(G.psi[k], F.psi[k]) < 1e-6, k


@pytest.mark.parametrize("free_top", [False, True])
def test_forward_seismograms_tiny(gpu_lib, free_top):
    p = tiny(n=(40, 22, 18), nt=150, W=5, seed=4, free_top=free_top, nrec=10)
    ref, _ = run_forward(p)
    prop = propagator_of(p)
    g, _ = prop.forward()
    got = g[0].double().cpu().numpy()
    assert np.abs(ref).max() > 0
    assert rel_l2(got, ref) < SEIS_TOL


def test_forward_is_deterministic_and_batch_equals_single(gpu_lib):
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    from tests.gpu_helpers import grid_of
    p = tiny(n=(36, 20, 16), nt=80, W=4, seed=6, nrec=7)
    grid = grid_of(p)
    rng = np.random.default_rng(0)
    srcs = [(10, 9, 8), (20, 12, 7), (25, 6, 10)]
    m6s = rng.uniform(-1, 1, (3, 6))
    nts = [80, 55, 71]
    W = np.stack([p.wavelet] * 3)
    dev = torch.device("cuda")
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device=dev)
    lam, mu, rho = f(p.lam), f(p.mu), f(p.rho)
    batch = Propagator(grid, lam, mu, rho, ShotBatch(grid, srcs, m6s, W, nts, [p.rec] * 3, p.cp))
    gb, _ = batch.forward()
    for s in range(3):
        single = Propagator(grid, lam, mu, rho, ShotBatch(grid, [srcs[s]], [m6s[s]], W[s:s + 1], [nts[s]],
                                                          [p.rec], p.cp))
        g1, _ = single.forward()
        assert torch.equal(gb[s, :, :nts[s]], g1[0, :, :nts[s]]), s
        assert not torch.any(gb[s, :, nts[s]:])


def _gradient_problem():
    return tiny(n=(30, 20, 18), nt=120, W=4, seed=12, nrec=9)


def test_exact_gradient_matches_oracle(gpu_lib):
    p = _gradient_problem()
    obs = run_forward(p, lam=p.lam * 1.02, rho=p.rho * 0.99)[0].astype(np.float32)
    hist = []
    g, _ = run_forward(p, history=hist)
    r, _ = O.misfit(g, obs)
    ref = O.exact_adjoint(p.medium(), p.cpml(), p.c, p.src, p.wavelet, p.rec, p.nt, r, hist)
    prop = propagator_of(p)
    J, gg, glam, gmu, grho, gm6 = prop.gradient(torch.as_tensor(obs[None], device="cuda"), K=11)
    assert rel_l2(gg[0].double().cpu().numpy(), g) < SEIS_TOL
    for name, got in (("lam", glam), ("mu", gmu), ("rho", grho)):
        e = rel_l2(got.double().cpu().numpy(), ref[name])
        assert e < GRAD_TOL, (name, e)
    assert rel_l2(gm6[0].cpu().numpy(), ref["m6"]) < GRAD_TOL


def test_exact_gradient_free_top_two_uneven_shots_match_oracle(gpu_lib):
    """Free surface on top; two shots in one batch with different sources (one
    three cells below the surface), MTs and record lengths (120 and 83 steps).
    Each shot's gathers and ∂J/∂M against its own oracle run, the volume
    gradients against the sum of the two oracle gradients, for the stored, the
    √-checkpointed (the short shot ends inside a segment) and the hybrid tape
    (the short shot ends inside the checkpointed part)."""
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    from tests.gpu_helpers import grid_of
    p = tiny(n=(30, 20, 18), nt=120, W=4, seed=31, nrec=9, free_top=True, src=(12, 9, 3))
    rng = np.random.default_rng(5)
    # the second MT is scaled so that both shots weigh alike in the summed gradients
    shots = [(p.src, p.m6, p.nt), ((19, 11, 12), 12 * rng.uniform(-1, 1, 6), 83)]
    true = O.Medium(p.lam * 1.02, p.mu, p.rho * 0.99)
    obs = np.zeros((2, 3, p.nt, len(p.rec)), np.float32)
    ref = {"lam": 0.0, "mu": 0.0, "rho": 0.0}
    g_ref, m6_ref, J_ref = [], [], 0.0
    for s, (src, m6, nt) in enumerate(shots):
        w = p.wavelet[:nt]
        obs[s, :, :nt] = O.forward(true, p.cpml(), p.c, src, m6, w, p.rec, nt)[0]
        hist = []
        g, _ = O.forward(p.medium(), p.cpml(), p.c, src, m6, w, p.rec, nt, history=hist)
        r, J = O.misfit(g, obs[s, :, :nt])
        a = O.exact_adjoint(p.medium(), p.cpml(), p.c, src, w, p.rec, nt, r, hist)
        for k in ref:
            ref[k] = ref[k] + a[k]
        g_ref.append(g); m6_ref.append(a["m6"]); J_ref += J
    grid = grid_of(p)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    batch = ShotBatch(grid, [s[0] for s in shots], [s[1] for s in shots], [p.wavelet] * 2,
                      [s[2] for s in shots], [p.rec] * 2, p.cp)
    prop = Propagator(grid, f(p.lam), f(p.mu), f(p.rho), batch)
    out = {}
    for strategy in ("store", "checkpoint", "hybrid"):
        kw = dict(m=30) if strategy == "hybrid" else dict(K=11)   # hybrid: steps 0-89 checkpointed
        J, gg, glam, gmu, grho, gm6 = prop.gradient(f(obs), strategy=strategy, **kw)
        assert prop.last_strategy == strategy
        out[strategy] = (glam, gmu, grho, gm6)
        assert abs(J - J_ref) <= 1e-4 * J_ref, (J, J_ref)
        for s, (_, _, nt) in enumerate(shots):
            assert rel_l2(gg[s, :, :nt].double().cpu().numpy(), g_ref[s]) < SEIS_TOL
            assert not gg[s, :, nt:].any()
            assert rel_l2(gm6[s].cpu().numpy(), m6_ref[s]) < GRAD_TOL, s
        for name, got in (("lam", glam), ("mu", gmu), ("rho", grho)):
            e = rel_l2(got.double().cpu().numpy(), ref[name])
            assert e < GRAD_TOL, (strategy, name, e)
    for other in ("checkpoint", "hybrid"):
        assert all(torch.equal(a, b) for a, b in zip(out["store"], out[other])), other


def test_exact_adjoint_given_the_oracle_residual(gpu_lib):
    """Adjoint precision alone: the device adjoint is driven by the oracle's
    residual, so forward FP32 error cannot leak in through the misfit."""
    p = _gradient_problem()
    obs = run_forward(p, mu=p.mu * 1.03)[0]
    hist = []
    g, _ = run_forward(p, history=hist)
    r, _ = O.misfit(g, obs)
    ref = O.exact_adjoint(p.medium(), p.cpml(), p.c, p.src, p.wavelet, p.rec, p.nt, r, hist)
    prop = propagator_of(p)
    _, snaps, K = prop.forward_checkpointed(K=13)
    glam, gmu, grho, gm6 = prop.gradient_from_snapshots(
        snaps, K, torch.as_tensor(r[None].astype(np.float32), device="cuda"))
    for name, got in (("lam", glam), ("mu", gmu), ("rho", grho)):
        e = rel_l2(got.double().cpu().numpy(), ref[name])
        assert e < 1e-5, (name, e)
    assert rel_l2(gm6[0].cpu().numpy(), ref["m6"]) < 1e-5


@pytest.mark.parametrize("fuse", [True, False])
def test_forward_wavefield_strategy_does_not_change_the_gradient(gpu_lib, fuse):
    """√nt checkpoints of any interval, hybrid tapes (the last m steps kept, the
    rest checkpointed) and keeping every state give identical bits, with the
    receivers sampled and injected inside the sweeps or by separate kernels."""
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    from tests.gpu_helpers import grid_of
    p = _gradient_problem()
    grid = grid_of(p)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    prop = Propagator(grid, f(p.lam), f(p.mu), f(p.rho),
                      ShotBatch(grid, [p.src], [p.m6], [p.wavelet], [p.nt], [p.rec], p.cp,
                                fuse_receivers=fuse))
    obs = torch.zeros((1, 3, p.nt, len(p.rec)), device="cuda")
    outs = [prop.gradient(obs, K=K, strategy="checkpoint") for K in (1, 7, p.nt)]
    # hybrid tapes: the last m steps kept, the first nt - m checkpointed
    for m in (11, p.nt // 2, p.nt - 3):
        outs.append(prop.gradient(obs, strategy="hybrid", m=m))
        assert prop.last_strategy == "hybrid"
    outs.append(prop.gradient(obs, strategy="store"))
    assert prop.last_strategy == "store"
    for o in outs[1:]:
        for a, b in zip(outs[0][2:], o[2:]):
            assert torch.equal(a, b)


def test_strategy_budget_counts_the_allocator_cache(gpu_lib):
    """Memory that torch's caching allocator holds but does not use counts as
    available: after a 4 GiB tensor is freed (and kept cached), the budget
    for keeping every forward state does not shrink, so the strategy does not
    fall back to a hybrid or checkpointed tape."""
    from elastic_fm_vp_vs_rho_amd.propagator import available_bytes
    dev = torch.device("cuda")
    torch.cuda.synchronize()
    a0 = available_bytes(dev)
    t = torch.empty(1 << 30, dtype=torch.float32, device=dev)
    del t
    free, _ = torch.cuda.mem_get_info(dev)
    a1 = available_bytes(dev)
    assert a1 >= a0 - (64 << 20), (a0, a1)
    assert a1 >= free
    p = _gradient_problem()
    prop = propagator_of(p)
    assert prop.choose_strategy() == "store"
    need = (p.nt + 1) * prop.state_bytes()
    assert prop.choose_strategy(budget=need - 1) == "hybrid"
    assert prop.choose_strategy(budget=8 * prop.state_bytes()) == "checkpoint"


def test_gpu_dot_product(gpu_lib):
    """⟨F m, r⟩ = ⟨m, Fᵀ r⟩ for the moment-tensor map, in FP32 on the device."""
    p = tiny(n=(34, 26, 22), nt=160, W=5, seed=15, nrec=12)
    prop = propagator_of(p)
    g, snaps, K = prop.forward_checkpointed()
    r = torch.randn(g.shape, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    _, _, _, gm6 = prop.gradient_from_snapshots(snaps, K, r)
    lhs = float(torch.sum(g.double() * r.double()))
    rhs = float(np.dot(p.m6, gm6[0].cpu().numpy()))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_graph_replay_equals_eager(gpu_lib):
    """The hipGraph path (one captured chunk replayed with a device step counter)
    gives the eager gathers bit for bit, on every replay, with eager work and
    synchronisations in between."""
    p = tiny(n=(36, 20, 18), nt=83, W=4, seed=17, nrec=6)
    prop = propagator_of(p)
    g_eager, st_eager = prop.forward()
    for chunk in (7, 20, 83):
        g_graph, st_graph = prop.forward(graph_chunk=chunk)
        torch.cuda.synchronize()
        assert torch.equal(g_graph, g_eager), chunk
        assert torch.equal(st_graph.fields, st_eager.fields), chunk
        g2, _ = prop.forward()                 # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(g2, g_eager)


def test_fused_receivers_equal_separate_kernels(gpu_lib):
    """Receiver sampling inside the velocity sweep and residual injection /
    ∂J/∂M inside the adjoint stress sweep (shots with a receiver bitmask) give
    the same bits as the separate per-step kernels, for a batch with uneven
    record lengths, on both forward-wavefield strategies."""
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    from tests.gpu_helpers import grid_of
    p = tiny(n=(30, 22, 20), nt=70, W=4, seed=41, nrec=1)
    grid = grid_of(p)
    rec = np.array([(5, 4, 3), (20, 10, 6), (12, 17, 15), (25, 5, 9), (8, 12, 2)])
    srcs = [(10, 9, 8), (20, 12, 7), (15, 10, 12)]
    m6s = np.random.default_rng(2).uniform(-1, 1, (3, 6))
    nts = [70, 41, 55]
    W = np.stack([p.wavelet] * 3)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    lam, mu, rho = f(p.lam), f(p.mu), f(p.rho)
    out = {}
    for fused in (True, False):
        sh = ShotBatch(grid, srcs, m6s, W, nts, [rec] * 3, p.cp, fuse_receivers=fused)
        assert (sh.recmask is not None) == fused
        prop = Propagator(grid, lam, mu, rho, sh)
        g, _ = prop.forward()
        gg, _ = prop.forward(graph_chunk=16)
        torch.cuda.synchronize()
        assert torch.equal(g, gg)
        obs = 0.5 * g.clone()
        res = [g]
        for strategy, K in (("checkpoint", 9), ("store", None)):
            J, _, glam, gmu, grho, gm6 = prop.gradient(obs, K=K, strategy=strategy)
            res += [glam, gmu, grho, gm6]
        out[fused] = res
    assert float(out[True][0].abs().max()) > 0
    for a, b in zip(out[True], out[False]):
        assert torch.equal(a, b)
