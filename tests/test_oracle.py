"""Pins for the FP64 oracle (CPU).

Each test is chosen so that a plausible transcription mistake (a dropped term, a
wrong sign, a shifted or transposed index, a mis-ordered ψ update) fails it:

* scalar-loop C twin vs vectorised NumPy, one full step from a random state;
* mirror and axis-permutation symmetry of an explosive source (pins the stagger
  and differencing directions, Q7);
* linearity in the moment tensor and zero-in/zero-out;
* CPML energy decay vs a reflecting (a=b=0) run;
* dot-product test of the exact adjoint (⟨F m, r⟩ = ⟨m, Fᵀ r⟩);
* central finite differences of the exact λ/μ/ρ/M gradients;
* reverse-time reconstruction with boundary restore reproduces the stored
  forward history in the interior (M:1929-1967).
"""
import numpy as np
import pytest

from oracle import elastic_fp64 as O
from oracle import twin
from tests.problems import rel_l2, run_forward, tiny

# a division by zero or an overflow inside the oracle is a bug, not a warning
pytestmark = pytest.mark.filterwarnings("error::RuntimeWarning")


def _random_state(shape, rng, scale=1.0):
    F = O.Fields(shape)
    for k in O.FIELDS:
        F.f[k][...] = scale * rng.standard_normal(shape)
    for k in O.PSI_V + O.PSI_T:
        F.psi[k][...] = 0.1 * scale * rng.standard_normal(shape)
    return F


def test_twin_matches_numpy_step():
    p = tiny(n=(11, 10, 9), W=3, seed=3)
    rng = np.random.default_rng(7)
    F1 = _random_state((p.ny, p.nz, p.nx), rng)
    F2 = F1.copy()
    med, cp = p.medium(), p.cpml()
    O.stress_sweep(F1, med, cp, p.c)
    O.add_source(F1, p.src, p.m6, 0.37)
    O.velocity_sweep(F1, med, cp, p.c)
    twin.step(F2, med.lam, med.mu, med.rho, cp, p.c, p.src, p.m6, 0.37)
    for k in O.FIELDS:
        assert rel_l2(F1.f[k], F2.f[k]) < 1e-13, k
    for k in O.PSI_V + O.PSI_T:
        assert rel_l2(F1.psi[k], F2.psi[k]) < 1e-13, k


def test_threaded_matches_serial():
    p = tiny(n=(12, 16, 10), nt=15, W=3, seed=4)
    g1, _ = run_forward(p, threads=1)
    g4, _ = run_forward(p, threads=4)
    assert np.array_equal(g1, g4)


def test_zero_source_gives_zero_field():
    p = tiny(nt=20)
    g, F = run_forward(p, m6=np.zeros(6))
    assert not np.any(g)
    assert all(not np.any(F.f[k]) for k in O.FIELDS)


def test_linearity_in_moment_tensor():
    p = tiny(nt=25, seed=5)
    basis = [run_forward(p, m6=np.eye(6)[i])[0] for i in range(6)]
    m = np.array([0.3, -1.2, 0.7, 0.25, -0.4, 0.9])
    g, _ = run_forward(p, m6=m)
    recon = sum(mi * bi for mi, bi in zip(m, basis))
    assert rel_l2(recon, g) < 1e-12


def test_explosive_source_symmetry():
    """Homogeneous cube, no CPML, isotropic source at the centre, few steps (the
    front has not reached the guard cells).  τ_ii live at integer points and
    v_x at x+½, so τ_xx(x) = τ_xx(2s-x) and v_x(x) = -v_x(2s-1-x); swapping the
    x and y axes maps v_x to v_y."""
    n, s = 21, 10
    p = tiny(n=(n, n, n), nt=3, W=0, seed=1, hetero=False, src=(s, s, s))
    cp = O.Cpml.zeros(n, n, n)
    F = O.Fields((n, n, n))
    med = p.medium()
    w = np.array([1.0, 0.6, -0.4])
    for it in range(3):
        O.stress_sweep(F, med, cp, p.c)
        O.add_source(F, (s, s, s), [1, 1, 1, 0, 0, 0], w[it])
        O.velocity_sweep(F, med, cp, p.c)
    f = F.f
    assert np.abs(f["vx"]).max() > 0
    X = np.arange(n)
    # x mirror (array axis 2)
    np.testing.assert_allclose(f["txx"], f["txx"][:, :, 2 * s - X], rtol=0, atol=1e-12 * np.abs(f["txx"]).max())
    np.testing.assert_allclose(f["vx"][:, :, :2 * s], -f["vx"][:, :, 2 * s - 1 - X[:2 * s]],
                               rtol=0, atol=1e-12 * np.abs(f["vx"]).max())
    # z mirror (axis 1) for v_z, y mirror (axis 0) for v_y
    np.testing.assert_allclose(f["vz"][:, :2 * s, :], -f["vz"][:, 2 * s - 1 - X[:2 * s], :],
                               rtol=0, atol=1e-12 * np.abs(f["vz"]).max())
    np.testing.assert_allclose(f["vy"][:2 * s], -f["vy"][2 * s - 1 - X[:2 * s]],
                               rtol=0, atol=1e-12 * np.abs(f["vy"]).max())
    # axis permutation x<->y: arrays are (y, z, x)
    np.testing.assert_allclose(f["vx"], np.transpose(f["vy"], (2, 1, 0)), rtol=0,
                               atol=1e-12 * np.abs(f["vx"]).max())
    np.testing.assert_allclose(f["txx"], np.transpose(f["tyy"], (2, 1, 0)), rtol=0,
                               atol=1e-12 * np.abs(f["txx"]).max())
    # x<->z: v_x(y, z, x) = v_z(y, x, z)
    np.testing.assert_allclose(f["vx"], np.transpose(f["vz"], (0, 2, 1)), rtol=0,
                               atol=1e-12 * np.abs(f["vx"]).max())


def _energy(F):
    return sum(float(np.sum(F.f[k] ** 2)) for k in ("vx", "vy", "vz"))


def test_cpml_absorbs():
    """With CPML on all faces the kinetic energy left after the pulse has passed
    is far below that of the same run with a=b=0 (Dirichlet walls reflect)."""
    n = 24
    p = tiny(n=(n, n, n), nt=1, W=6, seed=2, hetero=False, src=(12, 12, 12))
    med = p.medium()
    nt = 260
    w = O.ricker(0.12 / p.dt / 5, p.dt, nt)

    def run(cp):
        F = O.Fields((n, n, n))
        e = []
        for it in range(nt):
            O.stress_sweep(F, med, cp, p.c)
            O.add_source(F, p.src, [1, 1, 1, 0.3, 0, 0], w[it])
            O.velocity_sweep(F, med, cp, p.c)
            e.append(_energy(F))
        return np.array(e)

    e_pml = run(p.cpml())
    e_wall = run(O.Cpml.zeros(n, n, n))
    assert e_pml[-1] < 1e-2 * e_pml.max()
    assert e_wall[-1] > 0.2 * e_wall.max()


def test_cpml_profile_structure():
    W, N = 10, 64
    a, b = O.cpml_profile(10.0, 1.6e-3, 10.0, W, 3000.0, N)
    assert np.all(a[W:N - W] == 0) and np.all(b[W:N - W] == 0)
    # mirror symmetric, zero damping at the innermost cell of each side
    np.testing.assert_allclose(a, a[::-1], rtol=1e-6)
    assert a[W - 1] == 0 and a[N - W] == 0
    assert np.all((b[:W] > 0) & (b[:W] < 1))
    assert np.all(np.diff(a[:W]) > 0)                # |a| shrinks towards the interior
    cp = O.cpml_coefficients(W, 10.0, 1.6e-3, 10.0, 3000.0, N, N, N, free_top=True)
    assert not np.any(cp["az"][:W]) and not np.any(cp["bz"][:W]) and np.any(cp["az"][N - W:])


def test_ricker_shape():
    fq, dt = 10.0, 1e-3
    nt = 400
    w = O.ricker(fq, dt, nt)
    t0 = np.sqrt(2) / fq / dt
    assert abs(int(np.argmax(w)) - t0) <= 1
    # zero crossings at |t - t0| = 1/(sqrt(2) pi f)
    tz = t0 + 1.0 / (np.sqrt(2) * np.pi * fq) / dt
    i = int(tz)
    assert w[i] * w[i + 1] <= 0
    assert abs(float(np.sum(w))) < 1e-2 * float(np.sum(np.abs(w)))


def test_model_maps():
    rng = np.random.default_rng(0)
    vp = (2500 + 2000 * rng.random(1000)).astype(np.float32)
    vs = (vp / 1.73).astype(np.float32)
    rho = (2000 + 800 * rng.random(1000)).astype(np.float32)
    lam, mu = O.lame_from_velocity(vp, vs, rho)
    vp2, vs2 = O.velocity_from_lame(lam, mu, rho)
    assert rel_l2(vp2, vp) < 1e-6 and rel_l2(vs2, vs) < 1e-6
    # Nafe-Drake (Brocher 2005 polynomial, with the reference's 0.4712): 3 km/s -> ~2.23 g/cc
    assert abs(O.rho_nafe_drake(3000.0) - 2232.0) < 1.0


def _adjoint_problem(seed=11, nt=60):
    return tiny(n=(13, 12, 11), nt=nt, W=3, seed=seed, nrec=5)


def test_exact_adjoint_dot_product():
    p = _adjoint_problem()
    hist = []
    g, _ = run_forward(p, history=hist)
    r = np.random.default_rng(5).standard_normal(g.shape)
    gr = O.exact_adjoint(p.medium(), p.cpml(), p.c, p.src, p.wavelet, p.rec, p.nt, r, hist)
    lhs = float(np.sum(g * r))
    rhs = float(np.dot(p.m6, gr["m6"]))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


@pytest.mark.parametrize("which", ["lam", "mu", "rho", "m6"])
def test_exact_adjoint_finite_differences(which):
    p = _adjoint_problem(seed=12, nt=50)
    rng = np.random.default_rng(13)
    obs = run_forward(p, lam=p.lam * 1.02)[0] * 0.9
    hist = []
    g, _ = run_forward(p, history=hist)
    r, _ = O.misfit(g, obs)
    gr = O.exact_adjoint(p.medium(), p.cpml(), p.c, p.src, p.wavelet, p.rec, p.nt, r, hist)
    base = dict(lam=p.lam.astype(np.float64), mu=p.mu.astype(np.float64),
                rho=p.rho.astype(np.float64), m6=np.asarray(p.m6, np.float64))
    d = rng.standard_normal(base[which].shape) * (np.abs(base[which]).mean() if which != "m6" else 1.0)
    h = 1e-5

    def J(s):
        kw = dict(base)
        kw[which] = base[which] + s * h * d
        gg, _ = run_forward(p, m6=kw["m6"], lam=kw["lam"], mu=kw["mu"], rho=kw["rho"])
        return O.misfit(gg, obs)[1]

    fd = (J(1) - J(-1)) / (2 * h)
    an = float(np.sum(gr[which] * d))
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_velocity_chain_rule():
    p = _adjoint_problem(seed=14, nt=40)
    vp, vs, rho = (p.vp.astype(np.float64), p.vs.astype(np.float64), p.rho.astype(np.float64))
    obs = run_forward(p)[0] * 1.1

    def fwd(vp_, vs_, rho_, hist=None):
        mu = rho_ * vs_ * vs_
        lam = rho_ * vp_ * vp_ - 2.0 * mu
        return run_forward(p, lam=lam, mu=mu, rho=rho_, history=hist)[0]

    hist = []
    g = fwd(vp, vs, rho, hist)
    r, _ = O.misfit(g, obs)
    mu = rho * vs * vs
    med = O.Medium(rho * vp * vp - 2 * mu, mu, rho)
    gl = O.exact_adjoint(med, p.cpml(), p.c, p.src, p.wavelet, p.rec, p.nt, r, hist)
    gv = O.velocity_gradients(vp, vs, rho, gl)
    rng = np.random.default_rng(3)
    h = 1e-6   # the Vs gradient is a 14x cancellation of λ and μ terms: keep h small
    for name, base in (("vp", vp), ("vs", vs), ("rho", rho)):
        d = rng.standard_normal(base.shape) * base.mean()
        args = {"vp": vp, "vs": vs, "rho": rho}

        def J(s):
            a = dict(args)
            a[name] = base + s * h * d
            return O.misfit(fwd(a["vp"], a["vs"], a["rho"]), obs)[1]

        fd = (J(1) - J(-1)) / (2 * h)
        an = float(np.sum(gv[name] * d))
        assert abs(fd - an) <= 1e-6 * abs(an), (name, fd, an)


def test_reference_reconstruction_reproduces_interior_history():
    """M:1951-1967: reverse stepping with borders restored at index nt-it-2, CPML
    zones zeroed and the source removed at nt-it-1 gives back the forward v in
    the interior, for every step."""
    p = tiny(n=(20, 19, 18), nt=45, W=4, free_top=True, seed=21, src=(9, 9, 3))
    med, cp = p.medium(), p.cpml()
    hist, borders = [], {}
    g, F = O.forward(med, cp, p.c, p.src, p.m6, p.wavelet, p.rec, p.nt, history=hist,
                     borders=borders, W=p.W)
    W = p.W
    sl = (slice(W, p.ny - W), slice(0, p.nz - W), slice(W, p.nx - W))
    worst = 0.0
    for it in range(p.nt - 1):
        O.velocity_sweep(F, med, cp, -p.c)
        O.set_border(F, borders, p.nt - it - 2, W)
        O.zero_cpml_zone(F, W)
        O.stress_sweep(F, med, cp, -p.c)
        O.add_source(F, p.src, p.m6, -float(p.wavelet[p.nt - it - 1]))
        ref = hist[p.nt - 1 - it]
        for k in ("vx", "vy", "vz"):
            worst = max(worst, rel_l2(F.f[k][sl], ref.f[k][sl]))
    assert worst < 1e-9, worst
