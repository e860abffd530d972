"""torch.autograd path: hand-written adjoint behind ElasticModeling, checked
against the FP64 oracle gradient (tiny grid) and against central finite
differences on 20 random probe directions at the full 128³ × 2000-step
gradient workload (FP32 finite differences: tolerance 2e-3)."""
import numpy as np
import pytest
import torch

from oracle import elastic_fp64 as O
from tests.gpu_helpers import grid_of, shots_of
from tests.problems import rel_l2, run_forward, tiny

pytestmark = pytest.mark.gpu


def test_autograd_matches_oracle_velocity_gradients(gpu_lib):
    from elastic_fm_vp_vs_rho_amd.ops import elastic_modeling
    p = tiny(n=(30, 22, 20), nt=110, W=4, seed=21, nrec=7)
    grid = grid_of(p)
    shots = shots_of(p, grid)
    dev = "cuda"
    vp = torch.tensor(p.vp, device=dev, requires_grad=True)
    vs = torch.tensor(p.vs, device=dev, requires_grad=True)
    rho = torch.tensor(p.rho, device=dev, requires_grad=True)
    m6 = torch.tensor(np.asarray(p.m6, np.float32)[None], device=dev, requires_grad=True)
    obs_np = (0.7 * run_forward(p, lam=p.lam * 1.08)[0]).astype(np.float32)
    g = elastic_modeling(vp, vs, rho, m6, grid, shots, K=9)
    loss = 0.5 * ((g[0].double() - torch.tensor(obs_np, device=dev).double()) ** 2).sum()
    loss.backward()
    hist = []
    gr, _ = run_forward(p, history=hist)
    r, _ = O.misfit(gr, obs_np)
    ga = O.exact_adjoint(p.medium(), p.cpml(), p.c, p.src, p.wavelet, p.rec, p.nt, r, hist)
    gv = O.velocity_gradients(p.vp, p.vs, p.rho, ga)
    for name, t in (("vp", vp), ("vs", vs), ("rho", rho)):
        e = rel_l2(t.grad.double().cpu().numpy(), gv[name])
        assert e < 1e-4, (name, e)
    assert rel_l2(m6.grad[0].double().cpu().numpy(), ga["m6"]) < 1e-4


@pytest.mark.parametrize("K", [8, None])
def test_autograd_reused_shot_batch_and_flat_m6(gpu_lib, K):
    """Two differentiable calls on one shot batch with different moment tensors,
    both recorded before either backward, give the gradients of two separate
    batches: with checkpoints (K = 8) the backward re-runs forward segments after
    the second call changed the batch's M; with every state kept (K = None) the
    second call must not reuse the first call's states.  A (6,) moment tensor
    gets a (6,) gradient."""
    from elastic_fm_vp_vs_rho_amd.ops import elastic_modeling
    p = tiny(n=(26, 20, 18), nt=70, W=3, seed=22, nrec=5)
    grid = grid_of(p)
    dev = "cuda"
    obs = torch.tensor(run_forward(p, lam=p.lam * 1.05)[0].astype(np.float32), device=dev)
    ma = np.asarray(p.m6, np.float32)
    mb = np.roll(ma, 2) * np.float32(-0.7)

    def grads(shared):
        vp = torch.tensor(p.vp, device=dev, requires_grad=True)
        vs = torch.tensor(p.vs, device=dev, requires_grad=True)
        rho = torch.tensor(p.rho, device=dev, requires_grad=True)
        m1 = torch.tensor(ma, device=dev, requires_grad=True)
        m2 = torch.tensor(mb, device=dev, requires_grad=True)
        s1 = shots_of(p, grid)
        s2 = s1 if shared else shots_of(p, grid)
        g1 = elastic_modeling(vp, vs, rho, m1, grid, s1, K=K)
        g2 = elastic_modeling(vp, vs, rho, m2, grid, s2, K=K)
        loss = 0.5 * ((g1[0] - obs) ** 2).sum() + 0.5 * ((g2[0] - obs) ** 2).sum()
        loss.backward()
        assert m1.grad.shape == (6,)
        return [t.grad.double() for t in (vp, vs, rho, m1, m2)]

    for a, b in zip(grads(True), grads(False)):
        assert float((a - b).abs().max()) <= 1e-6 * float(b.abs().max())


def _smooth_probe(shape, gen, W, dev):
    d = torch.randn((1, 1) + shape, device=dev, generator=gen)
    k = torch.exp(-0.5 * (torch.arange(-6, 7, device=dev, dtype=torch.float32) / 3.0) ** 2)
    k = k / k.sum()
    for dim in (2, 3, 4):
        kk = k.view([1, 1] + [13 if i == dim else 1 for i in (2, 3, 4)])
        pad = [0, 0, 0, 0, 0, 0]
        pad[2 * (4 - dim)] = pad[2 * (4 - dim) + 1] = 6
        d = torch.nn.functional.conv3d(torch.nn.functional.pad(d, pad, mode="replicate"), kk)
    d = d[0, 0]
    m = torch.zeros_like(d)
    m[W:-W, W:-W, W:-W] = 1.0
    d = d * m
    return d / d.abs().max()


def test_c2_gradient_against_finite_differences(gpu_lib):
    """C2: 128³ three-layer model, anomaly in the true model, 2000 steps, one MT
    source, 64 receivers; 20 random directions perturbing Vp, Vs, ρ and M."""
    from elastic_fm_vp_vs_rho_amd import workloads
    from elastic_fm_vp_vs_rho_amd.ops import elastic_modeling
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, ShotBatch
    dev = torch.device("cuda")
    wl = workloads.c2()
    wt = workloads.c2(anomaly=True)
    grid = Grid(wl.nx, wl.ny, wl.nz, wl.dx, wl.dt, (wl.W,) * 6)
    shots = ShotBatch(grid, wl.src, wl.m6, wl.wavelets(), wl.nt, wl.rec, wl.cpml(), device=dev)
    f = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev)
    m6_0 = f(np.asarray(wl.m6, np.float32))
    with torch.no_grad():
        obs = elastic_modeling(f(wt.vp), f(wt.vs), f(wt.rho), m6_0, grid, shots).double()
    vp, vs, rho = f(wl.vp).requires_grad_(), f(wl.vs).requires_grad_(), f(wl.rho).requires_grad_()
    m6 = m6_0.clone().requires_grad_()
    J0 = 0.5 * ((elastic_modeling(vp, vs, rho, m6, grid, shots).double() - obs) ** 2).sum()
    J0.backward()
    grads = [vp.grad, vs.grad, rho.grad, m6.grad]
    base = [vp.detach(), vs.detach(), rho.detach(), m6.detach()]
    # probe amplitudes (m/s, m/s, kg/m³, MT units) balance central-difference
    # truncation against FP32 rounding of the 2000-step misfit
    scale = [1.25, 0.75, 2.0, 0.0025]
    gen = torch.Generator(device=dev).manual_seed(3)
    worst = 0.0

    def J(ps, h):
        with torch.no_grad():
            x = [b + h * s * d for b, s, d in zip(base, scale, ps)]
            g = elastic_modeling(x[0], x[1], x[2], x[3], grid, shots).double()
            return float(0.5 * ((g - obs) ** 2).sum())

    for _ in range(20):
        ps = [_smooth_probe(grid.shape, gen, wl.W + 2, dev) for _ in range(3)]
        ps.append(torch.randn((1, 6), device=dev, generator=gen))
        parts = [float((gg.double() * s * d.double()).sum()) for gg, s, d in zip(grads, scale, ps)]
        an = sum(parts)
        fd = (J(ps, 1.0) - J(ps, -1.0)) / 2.0
        # the four blocks differ by orders of magnitude and can cancel in a random
        # direction: measure the error against the blocks' total magnitude
        worst = max(worst, abs(fd - an) / sum(abs(p) for p in parts))
    assert worst < 2e-3, worst
