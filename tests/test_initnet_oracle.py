"""CPU checks of the FP64 restatement of the initializer network (the checker
of the HIP training kernels): BF16 rounding against torch's, im2col/col2im
adjointness, gradients against central finite differences on a tiny net, and
the host-side target encoding."""
import numpy as np
import pytest
import torch

from oracle import initnet_fp64 as R

TINY = R.Spec(cin=6, length=40, channels=5, ks=5, stride=2, pad=2, nblocks=3, hidden=7, nout=4)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-20, 20, 10000),
                        [0.0, -0.0, 1.0, -1.0, 3.0e38, -3.4e38, 1e-40, np.inf, -np.inf]]).astype(np.float32)
    # ties: exactly halfway between two BF16 values, both parities
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(R.bf16(x), want)
    assert np.isnan(R.bf16(np.array([np.nan], np.float32)))[0]


def test_im2col_col2im_adjoint():
    rng = np.random.default_rng(1)
    for (C, B, L, ks, s, p) in ((3, 2, 17, 5, 2, 2), (4, 3, 40, 9, 4, 4), (2, 1, 9, 3, 1, 0)):
        x = rng.standard_normal((C, B, L))
        col = R.im2col(x, ks, s, p)
        d = rng.standard_normal(col.shape)
        lhs = float(np.sum(col * d))
        rhs = float(np.sum(x * R.col2im(d, C, B, L, ks, s, p)))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
        # explicit definition of one column entry
        lo = (L + 2 * p - ks) // s + 1
        c, k, b, l = C - 1, ks // 2, B - 1, lo // 2
        pos = l * s + k - p
        assert col[c * ks + k, b * lo + l] == (x[c, b, pos] if 0 <= pos < L else 0.0)


def test_conv_forward_matches_direct_convolution():
    """The im2col GEMM equals the textbook strided cross-correlation."""
    rng = np.random.default_rng(2)
    p = R.init_params(TINY, seed=3)
    x = rng.standard_normal((TINY.cin, 2, TINY.length))
    col = R.im2col(x, TINY.ks, TINY.stride, TINY.pad)
    y = (p["W0"].astype(np.float64) @ col).reshape(TINY.channels, 2, -1)
    W = p["W0"].reshape(TINY.channels, TINY.cin, TINY.ks).astype(np.float64)
    xp = np.pad(x, ((0, 0), (0, 0), (TINY.pad, TINY.pad)))
    for co in range(TINY.channels):
        for b in range(2):
            for l in range(y.shape[2]):
                seg = xp[:, b, l * TINY.stride:l * TINY.stride + TINY.ks]
                assert abs(y[co, b, l] - np.sum(W[co] * seg)) < 1e-12


@pytest.mark.parametrize("emulate", [False, True])
def test_gradients_against_finite_differences(emulate):
    rng = np.random.default_rng(4)
    B = 3
    p = {k: v.astype(np.float64) for k, v in R.init_params(TINY, seed=5).items()}
    for k in p:
        if k[0] in "bc":
            p[k] = 0.1 * rng.standard_normal(p[k].shape)
    x = rng.standard_normal((TINY.cin, B, TINY.length))
    t = rng.standard_normal((TINY.nout, B))
    if emulate:
        # BF16 rounding is piecewise constant: check the plain network's FD
        # against the emulating backward only where rounding is the identity,
        # i.e. on inputs and weights already representable in BF16 and with
        # the FP64 path (activations rounded at the same places)
        x = R.bf16(x)
        p = {k: R.bf16(v) if k[0] in "WF" else v for k, v in p.items()}
    out, cache = R.forward(TINY, p, x, emulate=False)
    loss, g = R.backward(TINY, p, cache, t, emulate=False)

    def L(pp):
        o, _ = R.forward(TINY, pp, x, emulate=False)
        return float(np.mean((o - t) ** 2))

    assert abs(L(p) - loss) < 1e-14
    h = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for i in rng.choice(flat.size, min(6, flat.size), replace=False):
            pp = {kk: vv.copy() for kk, vv in p.items()}
            pp[k].reshape(-1)[i] += h
            lp = L(pp)
            pp[k].reshape(-1)[i] -= 2 * h
            lm = L(pp)
            fd = (lp - lm) / (2 * h)
            an = g[k].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-7, (k, i, fd, an)
    if emulate:
        # with every rounded operand already BF16, the emulating pass differs
        # from the plain one only by rounding activations and gradient operands
        _, ce = R.forward(TINY, p, x, emulate=True)
        le, ge = R.backward(TINY, p, ce, t, emulate=True)
        for k in p:
            rel = np.linalg.norm(ge[k] - g[k]) / max(np.linalg.norm(g[k]), 1e-30)
            assert rel < 0.05, (k, rel)


def test_adam_first_step_moves_by_lr():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    w1, m, v = R.adam(w, g, np.zeros(3), np.zeros(3), 1, lr=1e-2)
    # first bias-corrected Adam step is lr·sign(g) up to eps
    np.testing.assert_allclose(w1, w - 1e-2 * np.sign(g), rtol=0, atol=1e-7)


def test_pack_gathers_normalisation():
    rng = np.random.default_rng(6)
    g = rng.standard_normal((2, 3, 30, 4)).astype(np.float32)
    g[1] *= 1e-6
    x = R.pack_gathers(g, 25)
    assert x.shape == (2, 12, 25)
    for s in range(2):
        assert np.abs(x[s]).max() <= 1.0
        # channel c = comp*nrec + r, sample l
        assert x[s, 1 * 4 + 2, 7] == R.bf16(np.float32(g[s, 1, 7, 2]) * np.float32(1.0 / np.abs(g[s]).max()))


def test_targets_encode_decode_roundtrip():
    from elastic_fm_vp_vs_rho_amd import initnet_data as D
    from elastic_fm_vp_vs_rho_amd import workloads
    wl = workloads.c5(n=40, S=50, nt=16)
    lo, hi = wl.extra["src_range"]
    t = D.targets_of(wl.src, wl.m6, lo, hi)
    assert t.shape == (50, 9) and np.all(np.abs(t[:, 6:]) <= 1.0)
    m, pos = D.decode(t, lo, hi)
    np.testing.assert_array_equal(pos, wl.src)
    np.testing.assert_allclose(m, wl.m6, atol=1e-6)


def test_host_spec_matches_oracle_spec():
    from elastic_fm_vp_vs_rho_amd.initnet import NetSpec, he_init
    a, b = NetSpec(), R.Spec()
    assert a.convs() == b.convs() and a.shapes() == b.shapes() and a.flat == b.flat == 256
    pa, pb = he_init(a, 7), R.init_params(b, 7)
    for k in pa:
        np.testing.assert_array_equal(pa[k], pb[k])
