"""The initializer network's HIP kernels against the FP64 restatement
(oracle/initnet_fp64.py, BF16 rounding emulated where the device rounds).

Tolerances: a single GEMM 1e-5 (FP32 against FP64 accumulation of the same
BF16 operands); whole forward/backward passes 2e-3 relative L2 per tensor (FP32
accumulation plus BF16 rounding flips of activations that sit on a rounding
boundary); one Adam step 1e-6."""
import numpy as np
import pytest
import torch

from oracle import initnet_fp64 as R

pytestmark = pytest.mark.gpu


def _bf(a, dev="cuda"):
    return torch.as_tensor(np.asarray(a, np.float32), device=dev).to(torch.bfloat16)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("M,N,K,ta,tb", [(64, 64, 32, False, False), (70, 130, 45, False, False),
                                         (9, 5, 300, True, False), (33, 257, 16, False, True),
                                         (128, 17, 1000, True, True), (1, 1, 1, False, False),
                                         (100, 72, 40, False, False), (96, 24, 48, True, False),
                                         (65, 200, 2000, False, False)])
def test_gemm_against_fp64(gpu_lib, M, N, K, ta, tb):
    """C = A·B for row-major and transposed operand views, ragged edges, and the
    bias/ReLU epilogues (fixes the MFMA lane layout: a transposed or permuted
    tile fails by O(1)).  Shapes whose strides are multiples of 8 take the
    vectorised kernel (all four operand layouts occur), the others the
    element-wise one."""
    from elastic_fm_vp_vs_rho_amd.propagator import _ptr, _stream
    from elastic_fm_vp_vs_rho_amd import _capi
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = R.bf16(rng.standard_normal((M, K)))
    B = R.bf16(rng.standard_normal((K, N)))
    bias = rng.standard_normal(M).astype(np.float32)
    Ad = _bf(A.T.copy() if ta else A)
    Bd = _bf(B.T.copy() if tb else B)
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    want = A @ B
    for epi in (0, 1, 2):
        C = torch.full((M, N + 3), 7.0, device="cuda")
        Cb = torch.zeros((M, N + 3), dtype=torch.bfloat16, device="cuda")
        bd = torch.as_tensor(bias, device="cuda")
        _capi.check(gpu_lib.efm_net_gemm(M, N, K, _ptr(Ad), sam, sak, _ptr(Bd), sbk, sbn, _ptr(C), N + 3,
                                         _ptr(bd), _ptr(Cb), epi, _stream()), "gemm")
        torch.cuda.synchronize()
        got = C.cpu().numpy()
        w = want + (bias[:, None] if epi else 0.0)
        if epi == 1:
            w = np.maximum(w, 0.0)
        assert _rel(got[:, :N], w) < 1e-5, epi
        assert np.all(got[:, N:] == 7.0)                      # nothing written past N
        np.testing.assert_array_equal(Cb[:, :N].float().cpu().numpy(), R.bf16(got[:, :N]))


def test_im2col_col2im_flatten_kernels(gpu_lib):
    from elastic_fm_vp_vs_rho_amd.propagator import _ptr, _stream
    from elastic_fm_vp_vs_rho_amd import _capi
    rng = np.random.default_rng(3)
    C, B, L, ks, s, p = 5, 3, 37, 9, 4, 4
    lo = (L + 2 * p - ks) // s + 1
    x = R.bf16(rng.standard_normal((C, B, L)))
    xd = _bf(x)
    col = torch.zeros((C * ks, B * lo), dtype=torch.bfloat16, device="cuda")
    _capi.check(gpu_lib.efm_net_im2col(_ptr(xd), _ptr(col), C, B, L, ks, s, p, _stream()), "im2col")
    d = rng.standard_normal((C * ks, B * lo)).astype(np.float32)
    dx = torch.zeros((C, B, L), device="cuda")
    _capi.check(gpu_lib.efm_net_col2im(_ptr(torch.as_tensor(d, device="cuda")), _ptr(dx), C, B, L, ks, s, p,
                                       _stream()), "col2im")
    X = torch.zeros((C * L, B), dtype=torch.bfloat16, device="cuda")
    _capi.check(gpu_lib.efm_net_flatten(_ptr(xd), _ptr(X), C, B, L, _stream()), "flatten")
    dX = rng.standard_normal((C * L, B)).astype(np.float32)
    da = torch.zeros((C, B, L), device="cuda")
    _capi.check(gpu_lib.efm_net_unflatten(_ptr(torch.as_tensor(dX, device="cuda")), _ptr(da), C, B, L,
                                          _stream()), "unflatten")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(col.float().cpu().numpy(), R.im2col(x, ks, s, p))
    assert _rel(dx.cpu().numpy(), R.col2im(d.astype(np.float64), C, B, L, ks, s, p)) < 1e-6
    np.testing.assert_array_equal(X.float().cpu().numpy(), x.transpose(0, 2, 1).reshape(C * L, B))
    np.testing.assert_array_equal(da.cpu().numpy(), dX.reshape(C, L, B).transpose(0, 2, 1))


SMALL = dict(cin=12, length=200, channels=16, ks=5, stride=2, pad=2, nblocks=3, hidden=32, nout=9)


def test_forward_backward_against_oracle(gpu_lib):
    from elastic_fm_vp_vs_rho_amd.initnet import InitNet, NetSpec
    spec = NetSpec(**SMALL)
    ospec = R.Spec(**SMALL)
    rng = np.random.default_rng(11)
    B = 6
    p = R.init_params(ospec, seed=2)
    for k in p:
        if k[0] in "bc":
            p[k] = (0.05 * rng.standard_normal(p[k].shape)).astype(np.float32)
    net = InitNet(spec, params=p, device="cuda")
    x = R.bf16(rng.standard_normal((spec.cin, B, spec.length)))
    t = rng.standard_normal((spec.nout, B)).astype(np.float32)
    net.set_input(_bf(x))
    net._buffers(B)["t"].copy_(torch.as_tensor(t, device="cuda"))
    out = net.forward_buffers(B)
    loss = net.backward_buffers(B)
    torch.cuda.synchronize()
    ro, cache = R.forward(ospec, p, x, emulate=True)
    rl, rg = R.backward(ospec, p, cache, t.astype(np.float64), emulate=True)
    assert _rel(out.cpu().numpy(), ro) < 2e-3
    assert abs(float(loss) - rl) <= 2e-3 * rl
    for k, g in rg.items():
        e = _rel(net.g[k].cpu().numpy(), g)
        assert e < 2e-3, (k, e)
    # one Adam step against the restatement, from the device gradients
    w0 = {k: v.cpu().numpy().astype(np.float64) for k, v in net.w.items()}
    gd = {k: v.cpu().numpy().astype(np.float64) for k, v in net.g.items()}
    net.adam(lr=1e-3)
    torch.cuda.synchronize()
    for k in w0:
        w1, _, _ = R.adam(w0[k], gd[k], np.zeros_like(w0[k]), np.zeros_like(w0[k]), 1, lr=1e-3)
        np.testing.assert_allclose(net.w[k].cpu().numpy(), w1, rtol=0, atol=1e-6)
        np.testing.assert_array_equal(net.wb[k].float().cpu().numpy(), R.bf16(net.w[k].cpu().numpy()))


def test_graph_training_step_equals_eager(gpu_lib):
    """Training steps replayed from a captured hipGraph (batch assembly, forward,
    backward) give the same weights, bit for bit, as the same steps launched
    eagerly, over several replays with eager work in between that overwrites the
    graph's own buffers (a prediction at the same batch size), and for two batch
    sizes each with its own graph.  The loss (an atomic FP64 sum) agrees to
    1e-12."""
    from elastic_fm_vp_vs_rho_amd.initnet import InitNet, NetSpec
    spec = NetSpec(**SMALL)
    rng = np.random.default_rng(21)
    N = 40
    data = _bf(rng.standard_normal((N, spec.cin, spec.length)))
    tg = torch.as_tensor(rng.standard_normal((N, 9)).astype(np.float32), device="cuda")
    eager = InitNet(spec, seed=4, device="cuda")
    graph = InitNet(spec, seed=4, device="cuda")
    assert torch.equal(eager.flat_w, graph.flat_w)
    steps = [(6, 0), (6, 1), (4, 2), (6, 3), (4, 4), (6, 5), (6, 6)]
    for B, seed in steps:
        idx = torch.as_tensor(np.random.default_rng(seed).choice(N, B, replace=False).astype(np.int32),
                              device="cuda")
        le = eager.train_step(data, tg, idx, lr=1e-3)
        lg = graph.train_step(data, tg, idx, lr=1e-3, graph=True)
        torch.cuda.synchronize()
        assert abs(float(lg) - float(le)) <= 1e-12 * abs(float(le)), (B, seed)
        assert torch.equal(graph.flat_w, eager.flat_w), (B, seed)
        assert torch.equal(graph.flat_wb, eager.flat_wb), (B, seed)
        # eager work between replays, on the graph's own B-sized buffers
        p_g = graph.predict(data, idx[:B].flip(0), batch=B)
        p_e = eager.predict(data, idx[:B].flip(0), batch=B)
        torch.cuda.synchronize()
        assert torch.equal(p_g, p_e)
    assert len(graph._graphs) == 2
    assert graph.t == eager.t == len(steps)


def test_pack_and_gather_batch(gpu_lib):
    from elastic_fm_vp_vs_rho_amd.initnet import InitNet, NetSpec, pack_gathers
    rng = np.random.default_rng(5)
    S, nrec, nt, L = 5, 4, 40, 32
    g = rng.standard_normal((S, 3, nt, nrec)).astype(np.float32)
    g[2] *= 1e-5
    data = torch.zeros((S, 3 * nrec, L), dtype=torch.bfloat16, device="cuda")
    pack_gathers(torch.as_tensor(g, device="cuda"), L, data, 0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(data.float().cpu().numpy(), R.pack_gathers(g, L))
    spec = NetSpec(cin=3 * nrec, length=L, channels=8, ks=3, stride=2, pad=1, nblocks=2, hidden=8)
    net = InitNet(spec, device="cuda")
    tg = torch.as_tensor(rng.standard_normal((S, 9)).astype(np.float32), device="cuda")
    idx = torch.tensor([3, 0, 4], dtype=torch.int32, device="cuda")
    B = net.load_batch(data, tg, idx)
    b = net._buffers(B)
    torch.cuda.synchronize()
    assert torch.equal(b["x"], data[idx.long()].permute(1, 0, 2))
    assert torch.equal(b["t"], tg[idx.long()].t())


def test_training_on_propagator_data_learns(gpu_lib):
    """Small C5-like problem (40³, 768 shots, 256 steps, 16 receivers): the
    network trained by the HIP kernels predicts held-out source locations
    better than the training mean (depth by far), and all nine outputs
    together better than it.  (The full 96³ run learns every output; this
    size only checks that training on propagator data works.)"""
    from elastic_fm_vp_vs_rho_amd import initnet_data as D
    from elastic_fm_vp_vs_rho_amd import workloads
    from elastic_fm_vp_vs_rho_amd.initnet import InitNet, NetSpec
    wl = workloads.c5(n=40, S=768, nt=256, nrec=16)
    c5 = D.C5(wl, "cuda")
    data, tg, _ = c5.generate(np.arange(768), 256, batch=256)
    spec = NetSpec(cin=48, length=256, channels=32, ks=9, stride=4, pad=4, nblocks=3, hidden=64)
    net = InitNet(spec, seed=1, device="cuda")
    va = np.arange(128)
    hist = D.train(net, data, tg, np.arange(128, 768), va, epochs=30, batch=32, lr=2e-3)
    assert all(np.isfinite(h).all() for h in hist)
    assert hist[-1][0] < 0.5 * hist[0][0]
    vi = torch.as_tensor(va.astype(np.int32), device="cuda")
    pred = net.predict(data, vi)
    t_va, t_tr = tg[vi.long()], tg[128:]
    err = ((pred - t_va) ** 2).mean(0)
    base = ((t_va - t_tr.mean(0)) ** 2).mean(0)
    assert float(err[6:].mean()) < 0.75 * float(base[6:].mean()), (err, base)  # location
    assert float(err[8]) < 0.5 * float(base[8]), (err, base)                    # depth
    assert float(err.mean()) < float(base.mean()), (err, base)                  # all outputs
    # the joint Vp/Vs/ρ/MT A/B on two validation shots: both runs start from the
    # same smoothed model, every accepted L-BFGS step lowers the misfit
    jab = D.joint_ab(c5, net, data[:2], np.arange(2), iters=3, tols=(0.5,))
    for who in ("plain", "network"):
        J = jab["J"][who]
        assert 2 <= len(J) <= 4 and np.isfinite(J).all()
        assert all(b < a for a, b in zip(J, J[1:])), J
    assert set(jab["iters_to_tol"]) == {"plain@0.5", "network@0.5"}


def test_network_start_needs_fewer_lbfgs_iterations(gpu_lib):
    """The A/B the initializer exists for, at a reduced size (40³, 2048
    training shots of 256 steps, 16 receivers): on 8 held-out shots at their
    catalogue locations, L-BFGS on M from the network's proposal reaches
    1e-2·J0 in fewer iterations (median) than from the reference's fixed start
    Mmod = (0.5, 0, 0.5, 0, 0.5, 0) (M:2777), J0 being the plain start's misfit,
    and is never worse than it on more than two shots."""
    from elastic_fm_vp_vs_rho_amd import initnet_data as D
    from elastic_fm_vp_vs_rho_amd import workloads
    from elastic_fm_vp_vs_rho_amd.initnet import InitNet, NetSpec
    S, held = 2048, 8
    wl = workloads.c5(n=40, S=S + held, nt=256, nrec=16)
    c5 = D.C5(wl, "cuda")
    data, tg, _ = c5.generate(np.arange(S + held), 256, batch=256)
    spec = NetSpec(cin=48, length=256, channels=32, ks=9, stride=4, pad=4, nblocks=3, hidden=64)
    net = InitNet(spec, seed=1, device="cuda")
    D.train(net, data, tg, np.arange(128, S), np.arange(128), epochs=40, batch=64, lr=2e-3)
    heldout = np.arange(S, S + held)
    ab = D.ab_test(c5, net, data[S:], heldout, iters=12, tols=(1e-2,))
    it = lambda who: np.array([r["iters_to_tol"][f"catalogue/{who}@0.01"] or 13 for r in ab])
    plain, network = it("plain"), it("network")
    print(f"iterations to 1e-2 J0, plain {plain.tolist()} network {network.tolist()}")
    assert np.median(network) < np.median(plain), (plain, network)
    assert int(np.sum(network > plain)) <= 2, (plain, network)
