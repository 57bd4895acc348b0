"""GPU parity of the coarse-grid correction (rgbdflow_params.coarse_levels,
not in the reference): level solves, whole pairs with boundary maps made on
the device, batches of uneven sizes (different coarse depths in one step)
and graph replays, each bit for bit equal to the oracle run with the same
correction, with equal per-level iteration counts.  The oracle's coarse
operator, restriction and prolongation are pinned in
tests/test_oracle_pins_coarse.py."""
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


def _variant(R, name):
    return R.RELAX_TILED if name == "tiled" else R.RELAX_PLAIN


@pytest.mark.parametrize("variant", ["plain", "tiled"])
@pytest.mark.parametrize("levels", [1, 3, 8])
@pytest.mark.parametrize("exits", [True, False])
def test_level_solve_matches_oracle(ctx, R, variant, levels, exits):
    O = oracle()
    W, H = 96, 72
    li, (u, v, w), _ = level_inputs(seed=9, W=W, H=H)
    kw = dict(n_outer=4, n_sor=60)
    po = O.make_params(W, H, cg_levels=levels, cg_sweeps=2, cg_at=9, disable_inner_exit=not exits,
                       disable_outer_exit=not exits, **kw)
    pg = R.make_params(W, H, coarse_levels=levels, coarse_sweeps=2, coarse_at=9, relax_variant=_variant(R, variant),
                       inner_exit=exits, outer_exit=exits, **kw)
    print(f"level solve {W}x{H} {variant} levels {levels} exits {exits}", flush=True)
    ua, va, wa, sa = O.level_solve(li, po, True, u, v, w)
    ub, vb, wb, sb = ctx.solve_level(li, pg, True, u, v, w)
    assert sb.outer_exec[0] == sa.outer_exec and sb.sweeps_exec[0] == sa.sweeps_exec
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


@pytest.mark.parametrize("W,H", [(75, 53), (33, 21), (130, 67)])
@pytest.mark.parametrize("sweeps,at", [(1, 0), (3, 23), (2, 14)])
def test_level_solve_odd_sizes(ctx, R, W, H, sweeps, at):
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=21, W=W, H=H)
    kw = dict(n_outer=3, n_sor=50)
    po = O.make_params(W, H, cg_levels=8, cg_sweeps=sweeps, cg_at=at, **kw)
    pg = R.make_params(W, H, coarse_levels=8, coarse_sweeps=sweeps, coarse_at=at, relax_variant=R.RELAX_TILED, **kw)
    print(f"level solve {W}x{H} sweeps {sweeps} at {at}", flush=True)
    ua, va, wa, sa = O.level_solve(li, po, True, u, v, w)
    ub, vb, wb, sb = ctx.solve_level(li, pg, True, u, v, w)
    assert sb.sweeps_exec[0] == sa.sweeps_exec
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


@pytest.mark.parametrize("W,H", [(96, 72), (131, 67)])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_level_solve_streamed_shapes(ctx, R, W, H, storage):
    """Every streamed workgroup shape (the autotuner's usual choice at the
    finest levels) with launches of 1, 2 and its most iterations, the
    V-cycle splitting the launches of each inner iteration."""
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=41, W=W, H=H)
    kw = dict(n_outer=3, n_sor=50)
    okw = dict(record_storage=1) if storage == "f16" else {}
    gkw = dict(storage=R.STORE_F16) if storage == "f16" else {}
    po = O.make_params(W, H, cg_levels=8, cg_sweeps=2, cg_at=7, **okw, **kw)
    ua, va, wa, sa = O.level_solve(li, po, True, u, v, w)
    shapes = [(i, sh["streamed"]) for i, sh in enumerate(R.relax_shapes()) if sh["streamed"] > 0]
    assert shapes
    for cfg, tr in shapes:
        for ti in sorted({1, 2, tr}):
            print(f"level solve {W}x{H} {storage} streamed shape {cfg} {ti} per launch", flush=True)
            pg = R.make_params(W, H, coarse_levels=8, coarse_sweeps=2, coarse_at=7, relax_variant=R.RELAX_TILED,
                               tile_config=cfg, tile_iters=ti, **gkw, **kw)
            ub, vb, wb, sb = ctx.solve_level(li, pg, True, u, v, w)
            assert sb.sweeps_exec[0] == sa.sweeps_exec, (cfg, ti)
            assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb), (cfg, ti)


@pytest.mark.parametrize("rule", ["gs", "block", "f16", "bf16"])
def test_level_solve_rules_and_storage(ctx, R, rule):
    """The coarse levels relax with the level's own rule; with 16-bit records
    the coarse operator is made from the decoded fine records (f32)."""
    O = oracle()
    W, H = 96, 72
    li, (u, v, w), _ = level_inputs(seed=13, W=W, H=H)
    kw = dict(n_outer=3, n_sor=50)
    okw, gkw = {}, {}
    if rule in ("gs", "block"):
        okw = dict(omega=1.8, pixel_gs=1 if rule == "gs" else 2)
        gkw = dict(omega=1.8, pixel_update=R.PIXEL_GS if rule == "gs" else R.PIXEL_BLOCK)
    else:
        okw = dict(record_storage=1 if rule == "f16" else 2)
        gkw = dict(storage=R.STORE_F16 if rule == "f16" else R.STORE_BF16)
    po = O.make_params(W, H, cg_levels=8, **okw, **kw)
    pg = R.make_params(W, H, coarse_levels=8, relax_variant=R.RELAX_TILED, **gkw, **kw)
    print(f"level solve {W}x{H} rule {rule}", flush=True)
    ua, va, wa, sa = O.level_solve(li, po, True, u, v, w)
    ub, vb, wb, sb = ctx.solve_level(li, pg, True, u, v, w)
    assert sb.sweeps_exec[0] == sa.sweeps_exec
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


@pytest.mark.parametrize("variant", ["plain", "tiled", "streamed"])
def test_pair_with_device_maps_matches_oracle(ctx, R, variant):
    """A whole 320x240 pair (10 x 30) with boundary maps made on the device,
    as sceneFlowMultiscale.cpp:116 makes them."""
    O = oracle()
    cfg = WORKLOADS["320x240_10x30"]
    W, H = cfg["W"], cfg["H"]
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 17, cfg["segments"], cfg["f"], cfg["cx"], cfg["cy"])
    kw = dict(f=cfg["f"], cx=cfg["cx"], cy=cfg["cy"], n_outer=cfg["n_outer"], n_sor=cfg["n_sor"],
              min_width=cfg["min_width"])
    po = O.make_params(W, H, cg_levels=8, **kw)
    if variant == "streamed":
        # the largest streamed shape at every level
        shape = max((sh["streamed"], i) for i, sh in enumerate(R.relax_shapes()) if sh["streamed"] > 0)[1]
        kw = dict(kw, tile_config=shape)
    pg = R.make_params(W, H, coarse_levels=8, relax_variant=_variant(R, variant), **kw)
    print(f"pair {W}x{H} device maps {variant}", flush=True)
    ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, None, po)
    ub, vb, wb, sb = ctx.multiscale(pg, im1, d1, im2, d2, sigma, None)
    L = sa.n_levels
    assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
    assert [sb.outer_exec[k] for k in range(L)] == [sa.outer_exec[k] for k in range(L)]
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


def test_uneven_batch_and_graph_replays(ctx, R):
    """Pairs whose levels have different coarse depths share each step's
    launches; with params.graph the schedule is captured and replayed, and
    every replay equals the oracle."""
    O = oracle()
    sizes = [(200, 150, 20), (160, 120, 20), (64, 48, 12), (130, 66, 20)]
    cases = []
    for i, (W, H, mw) in enumerate(sizes):
        pair, im1, d1, im2, d2, sigma = make_case(W, H, 70 + i, 4)
        hpb, _ = hpb_for(pair, W, H, 0.5, mw)
        cases.append(dict(W=W, H=H, im1=im1, dm1=d1, im2=im2, dm2=d2, sigma1=sigma, hpb=hpb, min_width=mw))
    kw = dict(n_outer=3, n_sor=30)
    ref = []
    for c in cases:
        po = O.make_params(c["W"], c["H"], cg_levels=4, cg_sweeps=2, cg_at=4, min_width=c["min_width"], **kw)
        ref.append(O.multiscale(c["im1"], c["dm1"], c["im2"], c["dm2"], c["sigma1"], c["hpb"], po))
    W0, H0 = sizes[0][:2]
    p = R.make_params(W0, H0, coarse_levels=4, coarse_sweeps=2, coarse_at=4, min_width=sizes[0][2],
                      relax_variant=R.RELAX_TILED, graph=True, **kw)
    pairs = []
    for c in cases:
        q = dict(c)
        q["cam"] = R.make_params(c["W"], c["H"]).cam
        pairs.append(q)
    ctx.batch_load(p, pairs)
    for run in range(3):
        print(f"uneven batch run {run}", flush=True)
        stats = ctx.batch_run(p, stats=True)
        for i, c in enumerate(cases):
            ua, va, wa, sa = ref[i]
            ub, vb, wb = ctx.batch_read(i)
            L = sa.n_levels
            assert [stats[i].sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)], (run, i)
            assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb), (run, i)


@pytest.mark.parametrize("case", ["two_inner", "residual_exit"])
def test_level_solve_inner_iterations_and_residual_exit(ctx, R, case):
    """Several inner iterations (the coarse operators made again for each
    assembled system) and the FP64 residual exit (its checks between the
    V-cycles)."""
    O = oracle()
    W, H = 96, 72
    li, (u, v, w), _ = level_inputs(seed=31, W=W, H=H)
    if case == "two_inner":
        okw, gkw = dict(n_outer=3, n_inner=2, n_sor=30), dict(n_outer=3, n_inner=2, n_sor=30)
    else:
        okw = dict(n_outer=3, n_sor=75, residual_exit=True, residual_tol=0.05)
        gkw = dict(n_outer=3, n_sor=75, inner_exit=R.EXIT_RESIDUAL, residual_tol=0.05)
    po = O.make_params(W, H, cg_levels=8, **okw)
    pg = R.make_params(W, H, coarse_levels=8, relax_variant=R.RELAX_TILED, **gkw)
    print(f"level solve {W}x{H} {case}", flush=True)
    ua, va, wa, sa = O.level_solve(li, po, True, u, v, w)
    ub, vb, wb, sb = ctx.solve_level(li, pg, True, u, v, w)
    assert sb.sweeps_exec[0] == sa.sweeps_exec
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


def test_with_init_matches_oracle(ctx, R):
    """sceneFlowWithInit (one full-resolution level from an estimate) with
    the correction."""
    O = oracle()
    W, H = 128, 96
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 14, 4)
    hpb, _ = hpb_for(pair, W, H, 0.5, 16)
    hpb0 = hpb[: W * H * 2]
    est = pair["flow"].astype(np.float64) * 0.9
    po = O.make_params(W, H, n_outer=3, n_sor=40, cg_levels=8)
    pg = R.make_params(W, H, n_outer=3, n_sor=40, coarse_levels=8, relax_variant=R.RELAX_TILED)
    print(f"with_init {W}x{H}", flush=True)
    ua, va, wa, sa = O.with_init(im1, d1, im2, d2, sigma, hpb0, est, po)
    ub, vb, wb, sb = ctx.with_init(pg, im1, d1, im2, d2, est, sigma, hpb0)
    assert sb.sweeps_exec[0] == sa.sweeps_exec[0]
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


def test_bad_coarse_params_are_rejected(ctx, R):
    li, (u, v, w), _ = level_inputs(seed=9, W=32, H=24)
    for kw in (dict(coarse_levels=9), dict(coarse_levels=-1), dict(coarse_levels=2, coarse_sweeps=0),
               dict(coarse_levels=2, coarse_at=24), dict(coarse_levels=2, coarse_at=-1)):
        pg = R.make_params(32, 24, n_outer=1, n_sor=10, **kw)
        with pytest.raises(R.RgbdFlowError):
            ctx.solve_level(li, pg, True, u, v, w)
