"""Autotuner: every compiled tile reproduces the reference tile's output, and the
gradient computed with the tuned tiles equals the one with the default tiles."""
import pytest
import torch

from tests.gpu_helpers import propagator_of
from tests.problems import tiny

pytestmark = pytest.mark.gpu


def test_all_tiles_parity_clean_and_tuned_gradient_unchanged(gpu_lib):
    from elastic_fm_vp_vs_rho_amd import autotune
    p = tiny(n=(70, 24, 20), nt=60, W=4, seed=41, nrec=6)
    prop = propagator_of(p)
    obs = torch.zeros((1, 3, p.nt, len(p.rec)), device="cuda")
    for k in range(4):
        autotune.set_tile(k, autotune.REF_TILE)
    ref = prop.gradient(obs, K=8)
    rep = autotune.tune(prop, reps=3, use_cache=False)
    for name, r in rep.items():
        rejected = [t for t, v in r["ms"].items() if isinstance(v, str)]
        assert not rejected, (name, rejected)
        assert tuple(r["tile"]) == autotune.get_tile(autotune.KERNELS.index(name))
    got = prop.gradient(obs, K=8)
    for a, b in zip(ref[2:], got[2:]):
        assert float((a.double() - b.double()).abs().max()) <= 1e-6 * float(b.double().abs().max())
    for k in range(4):
        autotune.set_tile(k, autotune.DEFAULT_TILE)


def test_gradient_bits_do_not_depend_on_the_tile_mapping(gpu_lib):
    """Every block mapping computes each cell with the same operations, so the
    gradient of a two-shot batch (uneven record lengths) on a grid whose 66-cell
    interior rows do not fill 64-lane waves is bit-identical for the linear,
    XCD-aware and compact mappings (the compact one wraps waves across rows)."""
    import numpy as np
    from elastic_fm_vp_vs_rho_amd import autotune
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    from tests.gpu_helpers import grid_of
    p = tiny(n=(70, 24, 20), nt=60, W=4, seed=43, nrec=6)
    grid = grid_of(p)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    rng = np.random.default_rng(3)
    batch = ShotBatch(grid, [p.src, (20, 15, 9)], [p.m6, rng.uniform(-1, 1, 6)], [p.wavelet] * 2,
                      [60, 41], [p.rec] * 2, p.cp)
    prop = Propagator(grid, f(p.lam * 1.01), f(p.mu), f(p.rho), batch)
    obs = torch.zeros((2, 3, p.nt, len(p.rec)), device="cuda")
    out = {}
    try:
        for tile in (autotune.REF_TILE, (128, 1, 1, 8), (64, 1, 1, 16), (128, 1, 2, 16), (256, 1, 1, 16)):
            for k in range(4):
                autotune.set_tile(k, tile)
            J, g, *grads = prop.gradient(obs, K=8)
            out[tile] = (J, g, grads)
    finally:
        for k in range(4):
            autotune.set_tile(k, autotune.DEFAULT_TILE)
    J0, g0, grads0 = out[autotune.REF_TILE]
    assert J0 > 0 and float(grads0[0].abs().max()) > 0
    for tile, (J, g, grads) in out.items():
        assert torch.equal(g, g0), tile
        assert all(torch.equal(a, b) for a, b in zip(grads, grads0)), tile
        assert abs(J - J0) <= 1e-12 * J0, tile


def test_five_shot_batch_bits_do_not_depend_on_the_tile_mapping(gpu_lib):
    """The XCD-aware mappings deal each shot's tiles over the XCDs in turn; with
    five shots (a count that leaves padding blocks in every shot's deal), tiles
    of one, two and four rows and y chunks of two planes, the gathers and the
    gradient equal those of the linear mapping bit for bit."""
    import numpy as np
    from elastic_fm_vp_vs_rho_amd import autotune
    from elastic_fm_vp_vs_rho_amd.propagator import Propagator, ShotBatch
    from tests.gpu_helpers import grid_of
    p = tiny(n=(70, 23, 19), nt=50, W=4, seed=44, nrec=5)
    grid = grid_of(p)
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")
    rng = np.random.default_rng(5)
    srcs = [p.src, (20, 15, 9), (40, 8, 12), (12, 11, 7), (55, 16, 10)]
    batch = ShotBatch(grid, srcs, [rng.uniform(-1, 1, 6) for _ in srcs], [p.wavelet] * 5,
                      [50, 37, 50, 23, 44], [p.rec] * 5, p.cp)
    prop = Propagator(grid, f(p.lam * 1.01), f(p.mu), f(p.rho), batch)
    obs = torch.zeros((5, 3, p.nt, len(p.rec)), device="cuda")
    out = {}
    try:
        for tile in (autotune.REF_TILE, (64, 2, 1, 8), (64, 4, 2, 8), (128, 1, 2, 16), (64, 1, 1, 16)):
            for k in range(4):
                autotune.set_tile(k, tile)
            J, g, *grads = prop.gradient(obs, K=7)
            out[tile] = (J, g, grads)
    finally:
        for k in range(4):
            autotune.set_tile(k, autotune.DEFAULT_TILE)
    J0, g0, grads0 = out[autotune.REF_TILE]
    assert J0 > 0 and float(grads0[0].abs().max()) > 0
    for tile, (J, g, grads) in out.items():
        assert torch.equal(g, g0), tile
        assert all(torch.equal(a, b) for a, b in zip(grads, grads0)), tile
        assert abs(J - J0) <= 1e-12 * J0, tile
