"""The tiled (temporally blocked) relaxation against the oracle.

Every tile shape, workgroup shape and iterations-per-launch must give exactly
the values of the plain schedule: the interior of a tile is exact after T
iterations when the halo is 2T wide, and the per-pixel arithmetic is shared.
"""
import numpy as np
import pytest

from helpers import WORKLOADS, hpb_for, level_inputs, load_package, make_case, oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    return load_package()


@pytest.fixture(scope="module")
def ctx(R):
    c = R.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def system():
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=8, W=96, H=72)
    return O.level_system(li, O.make_params(96, 72), u, v, w)


@pytest.mark.parametrize("n_iter", [1, 3, 7, 25, 30])
@pytest.mark.parametrize("tile_iters", [1, 2, 4, 5])
def test_tiled_relax_matches_oracle(ctx, R, system, n_iter, tile_iters):
    O = oracle()
    x0 = np.zeros((3, 72, 96))
    a = O.relax(system, n_iter, x0)
    sched = R.make_params(96, 72, relax_variant=R.RELAX_TILED, tile_iters=tile_iters)
    b = ctx.relax(system, n_iter, x0, sched)
    assert np.array_equal(a, b), float(np.abs(a - b).max())


@pytest.mark.parametrize("cfg", range(5))
@pytest.mark.parametrize("tw,th", [(16, 8), (24, 12), (40, 16)])
def test_tiled_relax_every_config_and_tile(ctx, R, system, cfg, tw, th):
    O = oracle()
    rng = np.random.default_rng(cfg)
    x0 = rng.normal(scale=0.05, size=(3, 72, 96))
    a = O.relax(system, 10, x0)
    sched = R.make_params(96, 72, relax_variant=R.RELAX_TILED, tile_iters=3, tile_w=tw, tile_h=th, tile_config=cfg)
    b = ctx.relax(system, 10, x0, sched)
    assert np.array_equal(a, b), float(np.abs(a - b).max())


@pytest.mark.parametrize("W,H", [(40, 30), (33, 21), (80, 60)])
def test_tiled_relax_small_levels(ctx, R, W, H):
    """Whole-level tiles (no halo, 25 iterations per launch) and odd widths."""
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=10, W=W, H=H)
    coef = O.level_system(li, O.make_params(W, H), u, v, w)
    x0 = np.zeros((3, H, W))
    a = O.relax(coef, 50, x0)
    b = ctx.relax(coef, 50, x0, R.make_params(W, H, relax_variant=R.RELAX_TILED))
    assert np.array_equal(a, b), float(np.abs(a - b).max())


def test_tiled_level_solve_exits(ctx, R):
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=9, W=96, H=72)
    po = O.make_params(96, 72, n_outer=6, n_sor=60)
    pg = R.make_params(96, 72, n_outer=6, n_sor=60, relax_variant=R.RELAX_TILED)
    ua, va, wa, sa = O.level_solve(li, po, True, u, v, w)
    ub, vb, wb, sb = ctx.solve_level(li, pg, True, u, v, w)
    assert sb.outer_exec[0] == sa.outer_exec and sb.sweeps_exec[0] == sa.sweeps_exec
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


@pytest.mark.parametrize("name,seed", [("320x240_10x30", 11), ("640x480_15x50", 21)])
def test_tiled_multiscale_matches_oracle(ctx, R, name, seed):
    O = oracle()
    cfg = WORKLOADS[name]
    W, H = cfg["W"], cfg["H"]
    pair, im1, d1, im2, d2, sigma = make_case(W, H, seed, cfg["segments"], cfg["f"], cfg["cx"], cfg["cy"])
    hpb, lv = hpb_for(pair, W, H, 0.5, cfg["min_width"])
    kw = dict(f=cfg["f"], cx=cfg["cx"], cy=cfg["cy"], n_outer=cfg["n_outer"], n_sor=cfg["n_sor"],
              min_width=cfg["min_width"])
    ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb, O.make_params(W, H, **kw))
    ub, vb, wb, sb = ctx.multiscale(R.make_params(W, H, relax_variant=R.RELAX_TILED, **kw), im1, d1, im2, d2,
                                    sigma, hpb)
    L = sa.n_levels
    assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
    assert [sb.outer_exec[k] for k in range(L)] == [sa.outer_exec[k] for k in range(L)]
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


@pytest.mark.parametrize("tile_iters", [2, 5])
def test_tiled_relax_more_tiles_than_resident(ctx, R, tile_iters):
    """Regression: a workgroup's halo must read du from before the launch even
    when other workgroups of the launch have finished (640x480 has more tiles
    than run at once)."""
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=12, W=640, H=480)
    coef = O.level_system(li, O.make_params(640, 480), u, v, w)
    x0 = np.zeros((3, 480, 640))
    a = O.relax(coef, 20, x0)
    b = ctx.relax(coef, 20, x0, R.make_params(640, 480, relax_variant=R.RELAX_TILED, tile_iters=tile_iters))
    assert np.array_equal(a, b), float(np.abs(a - b).max())


def test_autotuned_schedule_is_exact(R):
    """Whatever schedule the autotuner picks, the solve is unchanged."""
    O = oracle()
    W, H = 320, 240
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 17, 4, 262.5, 159.5, 119.5)
    hpb, lv = hpb_for(pair, W, H, 0.5, 40)
    kw = dict(f=262.5, cx=159.5, cy=119.5, n_outer=6, n_sor=30, min_width=40)
    ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb, O.make_params(W, H, **kw))
    with R.Context(0) as c:
        for (w, h) in lv:
            ch = c.autotune_relax(w, h)
            assert ch["candidates"] > 0 and ch["us_per_iter"] > 0
        ub, vb, wb, sb = c.multiscale(R.make_params(W, H, relax_variant=R.RELAX_TILED, **kw), im1, d1, im2, d2,
                                      sigma, hpb)
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)
