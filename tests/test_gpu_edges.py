"""Edge cases on the GPU: no receivers at all, unused receiver slots, record
lengths of one step next to longer ones, the smallest grid with a single
interior cell and no absorbing layer, receivers off the interior, receivers
sharing a cell or sitting on the source, a different absorbing-layer width on
every face, and sources inside the layers of shots with their own profiles —
each against the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import elastic_fp64 as O
from tests.gpu_helpers import grid_of
from tests.problems import rel_l2, run_forward, tiny

pytestmark = pytest.mark.gpu


def _batch(p, srcs, m6s, nts, recs):
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    grid = grid_of(p)
    S = len(srcs)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    sh = ShotBatch(grid, srcs, m6s, np.stack([p.wavelet] * S), nts, recs, p.cp)
    return Propagator(grid, f(p.lam), f(p.mu), f(p.rho), sh)


def test_no_receivers_forward_and_gradient(gpu_lib):
    p = tiny(n=(24, 20, 18), nt=40, W=3, seed=31)
    prop = _batch(p, [p.src], [p.m6], [p.nt], np.zeros((1, 0, 3), np.int64))
    g, st = prop.forward()
    assert g.shape == (1, 3, p.nt, 0)
    assert float(st.fields.abs().max()) > 0                  # the source still radiates
    J, _, glam, gmu, grho, gm6 = prop.gradient(torch.zeros_like(g))
    assert J == 0.0
    for t in (glam, gmu, grho):
        assert not torch.any(t)
    assert not torch.any(gm6)


def test_unused_receiver_slots_stay_zero(gpu_lib):
    p = tiny(n=(24, 20, 18), nt=50, W=3, seed=32, nrec=5)
    rec = np.array(p.rec).copy()
    rec[[1, 3]] = -1                                          # slots 1 and 3 unused
    prop = _batch(p, [p.src], [p.m6], [p.nt], rec[None])
    g, _ = prop.forward()
    ref, _ = run_forward(p)
    got = g[0].double().cpu().numpy()
    assert not np.any(got[:, :, [1, 3]])
    keep = [0, 2, 4]
    assert rel_l2(got[:, :, keep], ref[:, :, keep]) < 1e-5


def test_one_step_records_next_to_longer_ones(gpu_lib):
    p = tiny(n=(26, 20, 18), nt=30, W=3, seed=33, nrec=6)
    rng = np.random.default_rng(1)
    m6s = rng.uniform(-1, 1, (3, 6))
    srcs = [p.src, (10, 9, 8), (14, 11, 9)]
    nts = [1, 30, 7]
    prop = _batch(p, srcs, m6s, nts, [p.rec] * 3)
    g, _ = prop.forward()
    for s in range(3):
        ref, _ = O.forward(p.medium(), p.cpml(), p.c, srcs[s], m6s[s], p.wavelet, p.rec, nts[s])
        got = g[s, :, :nts[s]].double().cpu().numpy()
        if np.abs(ref).max() == 0:                       # one step: nothing reaches the receivers yet
            assert not np.any(got), s
        else:
            assert rel_l2(got, ref) < 1e-5, s
        assert not torch.any(g[s, :, nts[s]:])
    # gradient of a batch whose shortest record is one step
    obs = torch.zeros_like(g)
    J, _, glam, gmu, grho, gm6 = prop.gradient(obs)
    assert np.isfinite(J) and torch.isfinite(glam).all() and torch.isfinite(gm6).all()


def test_smallest_grid_single_interior_cell(gpu_lib):
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, Propagator, ShotBatch
    p = tiny(n=(5, 5, 5), nt=12, W=0, seed=34, nrec=3, src=(2, 2, 2))
    grid = Grid(5, 5, 5, p.dx, p.dt, (0, 0, 0, 0, 0, 0))
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    sh = ShotBatch(grid, [p.src], [p.m6], [p.wavelet], [p.nt], [p.rec], p.cp)
    prop = Propagator(grid, f(p.lam), f(p.mu), f(p.rho), sh)
    g, _ = prop.forward()
    ref, _ = run_forward(p)
    assert np.abs(ref).max() > 0
    assert rel_l2(g[0].double().cpu().numpy(), ref) < 1e-5


def test_gradient_with_receivers_off_the_interior(gpu_lib):
    """Receivers on cells the sweeps never update (x = 1, y = 0, z = nz-1,
    x = nx-1) record zeros; their residuals are injected into the adjoint
    velocity but must not propagate (b̄ is zero off the interior): the gradient
    still matches the oracle to 1e-4."""
    import dataclasses
    from tests.gpu_helpers import propagator_of
    p = tiny(n=(30, 20, 18), nt=90, W=4, seed=15, nrec=5)
    extra = np.array([(1, 8, 7), (12, 0, 9), (10, 10, 17), (29, 5, 5)], np.int32)
    p = dataclasses.replace(p, rec=np.concatenate([p.rec, extra]))
    obs = run_forward(p, lam=p.lam * 1.02, rho=p.rho * 0.99)[0].astype(np.float32)
    obs[:, :, 5:] = np.random.default_rng(2).standard_normal(obs[:, :, 5:].shape) * np.abs(obs).max()
    hist = []
    g, _ = run_forward(p, history=hist)
    assert not np.any(g[:, :, 5:])
    r, _ = O.misfit(g, obs)
    ref = O.exact_adjoint(p.medium(), p.cpml(), p.c, p.src, p.wavelet, p.rec, p.nt, r, hist)
    prop = propagator_of(p)
    J, gg, glam, gmu, grho, gm6 = prop.gradient(torch.as_tensor(obs[None], device="cuda"), K=9)
    assert not torch.any(gg[0, :, :, 5:])
    for name, got in (("lam", glam), ("mu", gmu), ("rho", grho)):
        e = rel_l2(got.double().cpu().numpy(), ref[name])
        assert e < 1e-4, (name, e)
    assert rel_l2(gm6[0].cpu().numpy(), ref["m6"]) < 1e-4


@pytest.mark.parametrize("fuse", [True, False])
def test_different_cpml_width_on_every_face(gpu_lib, fuse):
    """A different absorbing-layer width on each face (x 5 / 3, y 3 / 5, z 0 /
    4; the C ABI's ψ slabs have one extent per face): seismograms and the exact
    gradient against the oracle on the same profiles, which are zero outside
    each face's layer (the oracle keeps full-volume ψ and full-length profiles,
    so the slab extents are the device's alone)."""
    from dataclasses import replace
    from elastic_fm_vp_vs_rho_amd.propagator import Grid, Propagator, ShotBatch
    p = tiny(n=(30, 26, 22), nt=110, W=5, seed=51, nrec=8, free_top=True)
    widths = (5, 3, 3, 5, 0, 4)
    cp = {k: np.array(v, np.float32) for k, v in p.cp.items()}
    for keys, n, (lo, hi) in ((("ax", "bx"), p.nx, widths[0:2]), (("ay", "by"), p.ny, widths[2:4]),
                              (("az", "bz"), p.nz, widths[4:6])):
        keep = np.zeros(n, bool)
        keep[:lo] = True
        keep[n - hi:] = True
        for k in keys:
            cp[k][..., ~keep] = 0
    q = replace(p, cp=cp)
    assert np.any(cp["ax"][..., 2:5]) and not np.any(cp["ax"][..., q.nx - 5:q.nx - 3])
    grid = Grid(q.nx, q.ny, q.nz, q.dx, q.dt, widths)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    prop = Propagator(grid, f(q.lam), f(q.mu), f(q.rho),
                      ShotBatch(grid, [q.src], [q.m6], [q.wavelet], [q.nt], [q.rec], q.cp,
                                fuse_receivers=fuse))
    obs = run_forward(q, lam=q.lam * 1.02, rho=q.rho * 0.99)[0].astype(np.float32)
    hist = []
    g_ref, _ = run_forward(q, history=hist)
    r, _ = O.misfit(g_ref, obs)
    ref = O.exact_adjoint(q.medium(), q.cpml(), q.c, q.src, q.wavelet, q.rec, q.nt, r, hist)
    print(f"widths {widths}, fuse {fuse}", flush=True)
    J, gg, glam, gmu, grho, gm6 = prop.gradient(f(obs[None]), K=13)
    assert rel_l2(gg[0].double().cpu().numpy(), g_ref) < 1e-5
    for name, got in (("lam", glam), ("mu", gmu), ("rho", grho)):
        assert rel_l2(got.double().cpu().numpy(), ref[name]) < 1e-4, name
    assert rel_l2(gm6[0].cpu().numpy(), ref["m6"]) < 1e-4


@pytest.mark.parametrize("fuse", [True, False])
def test_receivers_sharing_a_cell_and_on_the_source(gpu_lib, fuse):
    """Two receivers in one cell and one on the source cell: each slot records
    the same trace as the oracle and the exact gradient takes every slot's
    residual (the receiver mask marks a cell once; the sweeps loop over the
    slots that name it)."""
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    p = tiny(n=(26, 22, 20), nt=90, W=4, seed=61, nrec=5)
    rec = np.array(p.rec, copy=True)
    rec[1] = rec[0]                       # slots 0 and 1 share a cell
    rec[3] = p.src                        # slot 3 sits on the source
    p.rec = rec
    grid = grid_of(p)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    prop = Propagator(grid, f(p.lam), f(p.mu), f(p.rho),
                      ShotBatch(grid, [p.src], [p.m6], [p.wavelet], [p.nt], [rec], p.cp,
                                fuse_receivers=fuse))
    obs = run_forward(p, mu=p.mu * 1.03)[0].astype(np.float32)
    obs[:, :, 1] *= 0.5                   # the shared cell's two slots get different residuals
    hist = []
    g_ref, _ = run_forward(p, history=hist)
    r, _ = O.misfit(g_ref, obs)
    ref = O.exact_adjoint(p.medium(), p.cpml(), p.c, p.src, p.wavelet, rec, p.nt, r, hist)
    J, gg, glam, gmu, grho, gm6 = prop.gradient(f(obs[None]), K=9)
    got = gg[0].double().cpu().numpy()
    assert np.array_equal(got[:, :, 0], got[:, :, 1])
    assert rel_l2(got, g_ref) < 1e-5
    for name, t in (("lam", glam), ("mu", gmu), ("rho", grho)):
        assert rel_l2(t.double().cpu().numpy(), ref[name]) < 1e-4, name
    assert rel_l2(gm6[0].cpu().numpy(), ref["m6"]) < 1e-4


def test_sources_in_the_layers_with_per_shot_profiles(gpu_lib):
    """Two shots in one batch, each with its own CPML profiles (Ricker peaks 1.0
    and 1.6 × the base frequency: the layer's α follows the frequency) and its
    source inside the absorbing layers (an x-y edge, a y-z edge): each shot's
    gathers and ∂J/∂M against its own oracle run, the volume gradients against
    the sum of the two."""
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    p = tiny(n=(28, 24, 22), nt=100, W=5, seed=71, nrec=9)
    f0 = 0.12 / p.dt / 8
    vmax = float(p.vp.max())
    shots = []
    for k, (fq, src) in enumerate(((f0, (3, 4, 11)), (1.6 * f0, (14, 20, 18)))):
        cp = O.cpml_coefficients(p.W, fq, p.dt, p.dx, vmax, p.nx, p.ny, p.nz, free_top=False)
        shots.append(dict(cp=cp, src=src, wav=O.ricker(fq, p.dt, p.nt).astype(np.float32),
                          m6=np.random.default_rng(k).uniform(-1, 1, 6)))
    assert not np.allclose(shots[0]["cp"]["ax"], shots[1]["cp"]["ax"])
    grid = grid_of(p)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    prof = {k: np.stack([np.asarray(s["cp"][k], np.float32) for s in shots]) for k in shots[0]["cp"]}
    sh = ShotBatch(grid, [s["src"] for s in shots], [s["m6"] for s in shots], [s["wav"] for s in shots],
                   [p.nt] * 2, [p.rec] * 2, prof)
    prop = Propagator(grid, f(p.lam), f(p.mu), f(p.rho), sh)
    med_true = O.Medium(p.lam * 1.02, p.mu * 0.98, p.rho)
    obs = np.zeros((2, 3, p.nt, len(p.rec)), np.float32)
    refs, gref = [], {"lam": 0.0, "mu": 0.0, "rho": 0.0}
    for k, s in enumerate(shots):
        cpk = O.Cpml.from_dict(s["cp"])
        obs[k] = O.forward(med_true, cpk, p.c, s["src"], s["m6"], s["wav"], p.rec, p.nt)[0]
        hist = []
        g, _ = O.forward(p.medium(), cpk, p.c, s["src"], s["m6"], s["wav"], p.rec, p.nt, history=hist)
        r, _ = O.misfit(g, obs[k])
        a = O.exact_adjoint(p.medium(), cpk, p.c, s["src"], s["wav"], p.rec, p.nt, r, hist)
        refs.append((g, a["m6"]))
        for name in gref:
            gref[name] = gref[name] + a[name]
    J, gg, glam, gmu, grho, gm6 = prop.gradient(f(obs), K=10)
    for k, (g, m6) in enumerate(refs):
        assert rel_l2(gg[k].double().cpu().numpy(), g) < 1e-5, k
        assert rel_l2(gm6[k].cpu().numpy(), m6) < 1e-4, k
    for name, t in (("lam", glam), ("mu", gmu), ("rho", grho)):
        assert rel_l2(t.double().cpu().numpy(), gref[name]) < 1e-4, name
