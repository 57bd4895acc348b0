"""The streamed relaxation schedule against the oracle.

A streamed workgroup sweeps its tile from top to bottom through a ring of
region rows in LDS (colour pass p reaches row r at step r + 2p) instead of
holding the tile whole.  Every pixel update is still relax_update on the same
operands, so every streamed shape, tile and iterations-per-launch must give
exactly the values of the plain schedule (and of the oracle's relaxation,
src/sceneFlowSingleScaleJacobiCUDA.cu:382-470).
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
def streamed(R):
    """(tile_config, most iterations per launch) of every streamed shape."""
    s = [(i, sh["streamed"]) for i, sh in enumerate(R.relax_shapes()) if sh["streamed"] > 0]
    assert s, "no streamed workgroup shape"
    return s


@pytest.fixture(scope="module")
def system():
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=8, W=96, H=72)
    return O.level_system(li, O.make_params(96, 72), u, v, w)


@pytest.mark.parametrize("n_iter", [1, 3, 7, 25, 30])
@pytest.mark.parametrize("tw,th", [(96, 72), (48, 24), (40, 16), (96, 8), (16, 72), (24, 4)])
def test_stream_relax_matches_oracle(ctx, R, system, streamed, n_iter, tw, th):
    O = oracle()
    rng = np.random.default_rng(n_iter)
    x0 = rng.normal(scale=0.05, size=(3, 72, 96))
    a = O.relax(system, n_iter, x0)
    for cfg, tr in streamed:
        for ti in sorted({1, 2, tr}):
            print(f"shape {cfg} (ring {tr}) tile {tw}x{th} {ti} per launch, {n_iter} iterations", flush=True)
            sched = R.make_params(96, 72, relax_variant=R.RELAX_TILED, tile_iters=ti, tile_w=tw, tile_h=th,
                                  tile_config=cfg)
            b = ctx.relax(system, n_iter, x0, sched)
            assert np.array_equal(a, b), (cfg, ti, float(np.abs(a - b).max()))


@pytest.mark.parametrize("W,H", [(33, 21), (130, 67), (257, 45), (40, 30)])
def test_stream_relax_odd_levels(ctx, R, streamed, W, H):
    """Odd widths and heights, levels wider than one region row (several
    columns of tiles) and narrower ones, with the planner's tiles."""
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=10, W=W, H=H)
    coef = O.level_system(li, O.make_params(W, H), u, v, w)
    x0 = np.zeros((3, H, W))
    a = O.relax(coef, 50, x0)
    for cfg, tr in streamed:
        print(f"{W}x{H} shape {cfg}", flush=True)
        b = ctx.relax(coef, 50, x0, R.make_params(W, H, relax_variant=R.RELAX_TILED, tile_config=cfg))
        assert np.array_equal(a, b), (cfg, float(np.abs(a - b).max()))


def test_stream_relax_more_tiles_than_resident(ctx, R, streamed):
    """640x480 in short tiles: more tiles than workgroups run at once, so a
    tile's halo must read du from before the launch."""
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=12, W=640, H=480)
    coef = O.level_system(li, O.make_params(640, 480), u, v, w)
    x0 = np.zeros((3, 480, 640))
    a = O.relax(coef, 20, x0)
    for cfg, tr in streamed:
        tw = 128 - 4 * tr - 2 - 8
        print(f"640x480 shape {cfg} tile {tw}x24", flush=True)
        b = ctx.relax(coef, 20, x0, R.make_params(640, 480, relax_variant=R.RELAX_TILED, tile_config=cfg,
                                                  tile_iters=tr, tile_w=tw, tile_h=24))
        assert np.array_equal(a, b), (cfg, float(np.abs(a - b).max()))


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("mode", ["jacobi", "gs", "block"])
def test_stream_relax_storage_and_updates(ctx, R, system, streamed, storage, mode):
    """16-bit records and the point / block SOR updates: the streamed
    schedule gives the plain schedule's bits (the plain schedule is checked
    against the oracle by the precision tests)."""
    st = {"f32": R.STORE_F32, "f16": R.STORE_F16, "bf16": R.STORE_BF16}[storage]
    pu = {"jacobi": R.PIXEL_JACOBI, "gs": R.PIXEL_GS, "block": R.PIXEL_BLOCK}[mode]
    if pu == R.PIXEL_BLOCK and st != R.STORE_F32:
        st = R.STORE_F32  # the block update needs f32 records
    om = 0.7 if pu == R.PIXEL_JACOBI else 1.8
    x0 = np.zeros((3, 72, 96))
    a = ctx.relax(system, 30, x0, R.make_params(96, 72, relax_variant=R.RELAX_PLAIN, storage=st, omega=om,
                                                pixel_update=pu))
    for cfg, tr in streamed:
        print(f"{storage} {mode} shape {cfg}", flush=True)
        b = ctx.relax(system, 30, x0, R.make_params(96, 72, relax_variant=R.RELAX_TILED, tile_config=cfg,
                                                    tile_iters=tr, tile_w=40, tile_h=24, storage=st, omega=om,
                                                    pixel_update=pu))
        assert np.array_equal(a, b), (cfg, float(np.abs(a - b).max()))


def test_stream_multiscale_matches_oracle(ctx, R, streamed):
    """Whole 320x240 pairs (early exits at every check, level by level)
    with each streamed shape: the oracle's flow and iteration counts."""
    O = oracle()
    cfg0 = WORKLOADS["320x240_10x30"]
    W, H = cfg0["W"], cfg0["H"]
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 11, cfg0["segments"], cfg0["f"], cfg0["cx"], cfg0["cy"])
    hpb, lv = hpb_for(pair, W, H, 0.5, cfg0["min_width"])
    kw = dict(f=cfg0["f"], cx=cfg0["cx"], cy=cfg0["cy"], n_outer=cfg0["n_outer"], n_sor=cfg0["n_sor"],
              min_width=cfg0["min_width"])
    ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb, O.make_params(W, H, **kw))
    L = sa.n_levels
    for cfg, tr in streamed:
        print(f"320x240 pair, shape {cfg}", flush=True)
        ub, vb, wb, sb = ctx.multiscale(R.make_params(W, H, relax_variant=R.RELAX_TILED, tile_config=cfg, **kw),
                                        im1, d1, im2, d2, sigma, hpb)
        assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
        assert [sb.outer_exec[k] for k in range(L)] == [sa.outer_exec[k] for k in range(L)]
        assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)
