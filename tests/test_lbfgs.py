"""lbfgs.py parity: golden vectors from the reference module
(tests/golden/make_lbfgs_vectors.py) plus textbook properties of the two-loop
recursion (Nocedal & Wright, Alg. 7.4)."""
import os

import numpy as np
import pytest

from elastic_fm_vp_vs_rho_amd.lbfgs import L_BFGS, compute_sk_yk

GOLD = os.path.join(os.path.dirname(__file__), "golden", "lbfgs_vectors.npz")


@pytest.mark.parametrize("L", [2, 4, 5])
def test_matches_reference_vectors(L):
    d = np.load(GOLD)
    m, g, r = d[f"L{L}_m"], d[f"L{L}_g"], d[f"L{L}_r"]
    sk = np.zeros((L,) + m.shape[1:], np.float32)
    yk = np.zeros_like(sk)
    for it in range(m.shape[0]):
        compute_sk_yk(sk, yk, m[it], g[it], L, it)
        assert np.array_equal(sk, d[f"L{L}_sk"][it]), (L, it)
        assert np.array_equal(yk, d[f"L{L}_yk"][it]), (L, it)
        if it > 0:
            got = L_BFGS(sk, yk, g[it], L, it)
            np.testing.assert_allclose(got, r[it], rtol=2e-5, atol=2e-6 * np.abs(r[it]).max())


def test_last_slot_never_written():
    L = 4
    sk = np.zeros((L, 2, 2, 2)); yk = np.zeros_like(sk)
    rng = np.random.default_rng(0)
    for it in range(8):
        compute_sk_yk(sk, yk, rng.standard_normal((2, 2, 2)), rng.standard_normal((2, 2, 2)), L, it)
    assert not np.any(sk[L - 1]) and not np.any(yk[L - 1])


def test_secant_equation_and_1d_inverse():
    rng = np.random.default_rng(1)
    n = 6
    A = np.diag(rng.uniform(1.0, 4.0, n))
    L = 6
    sk = np.zeros((L, 1, 1, n)); yk = np.zeros_like(sk)
    m = rng.standard_normal(n)
    for it in range(4):
        compute_sk_yk(sk, yk, m.reshape(1, 1, n), (A @ m).reshape(1, 1, n), L, it)
        m = m - 0.3 * rng.standard_normal(n)
    # BFGS satisfies H y = s for the most recent pair
    np.testing.assert_allclose(L_BFGS(sk, yk, yk[1], L, 3), sk[1], rtol=1e-10, atol=1e-12)
    # one dimension: H = s/y exactly
    s1 = np.zeros((3, 1, 1, 1)); y1 = np.zeros_like(s1)
    compute_sk_yk(s1, y1, np.full((1, 1, 1), 2.0), np.full((1, 1, 1), 6.0), 3, 0)
    compute_sk_yk(s1, y1, np.full((1, 1, 1), 1.0), np.full((1, 1, 1), 3.0), 3, 1)
    np.testing.assert_allclose(L_BFGS(s1, y1, np.full((1, 1, 1), 5.0), 3, 1), 5.0 / 3.0)


def test_sign_invariance():
    """L_BFGS(s, -y, -g) = L_BFGS(s, y, g): the FWI driver relies on it when it
    feeds a descent-signed image as the gradient."""
    rng = np.random.default_rng(2)
    L = 5
    sk = rng.standard_normal((L, 2, 3, 4)); yk = rng.standard_normal((L, 2, 3, 4))
    g = rng.standard_normal((2, 3, 4))
    np.testing.assert_allclose(L_BFGS(sk, -yk, -g, L, 3), L_BFGS(sk, yk, g, L, 3), rtol=1e-12)
