"""GPU parity: every stage of the HIP path against the oracle on the same inputs.

The kernels evaluate the reference's expressions in the reference's order and
precision, compiled without FMA contraction, and the oracle does the same on
the CPU, so most stages are compared for exact (bit-level) equality.  Where a
libm transcendental enters (exp/log of the depth-noise inflation) a 1-ulp f32
difference is allowed.  Whole-pair results are held to the stated tolerance:
max |du|, |dv| <= 1e-3 px and |dw| <= 1e-5 m against the reference-precision
oracle at identical iteration counts; the distance to the FP64 oracle is
reported.
"""
import numpy as np
import pytest

from helpers import WORKLOADS, hpb_for, level_inputs, load_package, make_case, oracle

pytestmark = pytest.mark.gpu

TOL_UV = 1e-3   # px
TOL_W = 1e-5    # m


@pytest.fixture(scope="module")
def R():
    return load_package()


@pytest.fixture(scope="module")
def ctx(R):
    c = R.Context(0)
    yield c
    c.close()


def _maxdiff(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))


@pytest.mark.parametrize("shape,ratio,min_width", [((97, 73, 3), 0.5, 10), ((64, 48, 1), 0.75, 12),
                                                    ((160, 120, 3), 0.5, 20), ((1280, 720, 3), 0.5, 40),
                                                    ((33, 17, 1), 0.5, 3), ((131, 67, 1), 0.5, 4)])
def test_gaussian_pyramid_exact(ctx, shape, ratio, min_width):
    O = oracle()
    W, H, Cn = shape
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 1, size=(H, W, Cn) if Cn > 1 else (H, W))
    ref = O.gaussian_pyramid(img, ratio, min_width)
    got = ctx.gaussian_pyramid(img, ratio, min_width)
    assert len(ref) == len(got)
    for a, b in zip(ref, got):
        assert a.shape == b.shape
        assert np.array_equal(a, b), _maxdiff(a, b)


def test_im2feature_exact(ctx):
    O = oracle()
    rng = np.random.default_rng(2)
    bgr = rng.uniform(0, 1, size=(53, 71, 3))
    assert np.array_equal(O.im2feature(bgr), ctx.im2feature(bgr))


def test_upsample_exact(ctx):
    O = oracle()
    rng = np.random.default_rng(3)
    for (Ws, Hs, Wd, Hd) in [(40, 30, 80, 60), (60, 33, 120, 67), (21, 17, 43, 35)]:
        u, v, w = (rng.normal(size=(Hs, Ws)) for _ in range(3))
        got = ctx.upsample_flow(u, v, w, Wd, Hd, 2.0)
        exp = [O.resize_dims(u, Wd, Hd) * 2.0, O.resize_dims(v, Wd, Hd) * 2.0, O.resize_dims(w, Wd, Hd)]
        for a, b in zip(exp, got):
            assert np.array_equal(a, b), _maxdiff(a, b)


def test_warp_by_3d_flow_exact(ctx):
    O = oracle()
    rng = np.random.default_rng(4)
    H, W = 45, 61
    I1, I2 = rng.uniform(size=(H, W, 5)), rng.uniform(size=(H, W, 5))
    D1, D2 = rng.uniform(1, 3, size=(H, W)), rng.uniform(1, 3, size=(H, W))
    u, v, w = rng.normal(scale=3, size=(H, W)), rng.normal(scale=3, size=(H, W)), rng.normal(scale=.05, size=(H, W))
    a = O.warp_cpu(I1, D1, I2, D2, u, v, w)
    b = ctx.warp_by_3d_flow(I1, D1, I2, D2, u, v, w)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_depth_stdevs(ctx):
    O = oracle()
    pair, im1, d1, im2, d2, sigma = make_case(160, 120, 5, 4)
    d = d1.astype(np.float32)
    d[10:14, 20:30] = 0  # invalid pixels: sigma 1 and inflation around them
    a = O.depth_stdevs(d, 4)
    b = ctx.depth_stdevs(d, 4)
    ulp = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1, ulp.max()
    assert (ulp > 0).mean() < 1e-3


def test_depth_map_pyramid_exact(ctx):
    O = oracle()
    pair, im1, d1, im2, d2, sigma = make_case(160, 120, 6, 4)
    d = d1.copy()
    d[30:40, 50:70] = 0
    a = O.depth_map_pyramid(d, 0.5, 20)
    b = ctx.depth_map_pyramid(d, 0.5, 20)
    for x, y in zip(a, b):
        assert np.array_equal(x, y), _maxdiff(x, y)


@pytest.mark.parametrize("reg", [2, 0])
def test_level_system_exact(ctx, reg):
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=7, W=96, H=72, reg=reg)
    p = O.make_params(96, 72)
    a = O.level_system(li, p, u, v, w)
    b = ctx.level_system(li, p, u, v, w)
    for k in range(13):
        assert np.array_equal(a[k], b[k]), (k, _maxdiff(a[k], b[k]))


def test_relax_exact(ctx):
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=8, W=96, H=72)
    p = O.make_params(96, 72)
    coef = O.level_system(li, p, u, v, w)
    x0 = np.zeros((3, 72, 96))
    a = O.relax(coef, 30, x0)
    b = ctx.relax(coef, 30, x0)
    assert np.array_equal(a, b), _maxdiff(a, b)


@pytest.mark.parametrize("allow_early,exits", [(True, True), (False, False)])
def test_level_solve(ctx, R, allow_early, exits):
    O = oracle()
    li, (u, v, w), _ = level_inputs(seed=9, W=96, H=72)
    po = O.make_params(96, 72, n_outer=6, n_sor=60, disable_inner_exit=not exits, disable_outer_exit=not exits)
    pg = R.make_params(96, 72, n_outer=6, n_sor=60, inner_exit=exits, outer_exit=exits)
    ua, va, wa, sa = O.level_solve(li, po, allow_early, u, v, w)
    ub, vb, wb, sb = ctx.solve_level(li, pg, allow_early, u, v, w)
    assert sb.outer_exec[0] == sa.outer_exec and sb.sweeps_exec[0] == sa.sweeps_exec
    assert _maxdiff(ua, ub) <= TOL_UV and _maxdiff(va, vb) <= TOL_UV and _maxdiff(wa, wb) <= TOL_W
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


def _pair_parity(ctx, R, name, seed, exits=True):
    O = oracle()
    cfg = WORKLOADS[name]
    W, H = cfg["W"], cfg["H"]
    pair, im1, d1, im2, d2, sigma = make_case(W, H, seed, cfg["segments"], cfg["f"], cfg["cx"], cfg["cy"])
    hpb, lv = hpb_for(pair, W, H, 0.5, cfg["min_width"])
    kw = dict(f=cfg["f"], cx=cfg["cx"], cy=cfg["cy"], n_outer=cfg["n_outer"], n_sor=cfg["n_sor"],
              min_width=cfg["min_width"])
    po = O.make_params(W, H, **kw, disable_inner_exit=not exits, disable_outer_exit=not exits)
    pg = R.make_params(W, H, **kw, inner_exit=exits, outer_exit=exits)
    ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb, po)
    ub, vb, wb, sb = ctx.multiscale(pg, im1, d1, im2, d2, sigma, hpb)
    L = sa.n_levels
    assert sb.n_levels == L
    assert [sb.outer_exec[k] for k in range(L)] == [sa.outer_exec[k] for k in range(L)]
    assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
    # the outer early-exit statistic's closest approach to 4e-4 (the decisions above agree; the
    # margin says how far an f32 summation, as the reference's, would have to move it)
    for k in range(L):
        m = sb.stat_margin[k]
        if k < L - 1:   # the coarsest level has no early-exit test (sceneFlowMultiscale.cpp:175)
            assert 0 <= m <= abs(sb.last_stat[k] - 4e-4) / 4e-4 + 1e-12
        else:
            assert m == -1
    print(f"{name} seed {seed}: outer statistic margin per level {[round(sb.stat_margin[k], 4) for k in range(L)]}")
    du, dv, dw = _maxdiff(ua, ub), _maxdiff(va, vb), _maxdiff(wa, wb)
    assert du <= TOL_UV and dv <= TOL_UV and dw <= TOL_W, (du, dv, dw)
    # error against the FP64 oracle (reported; not a bound on the GPU path)
    po.precision = O.PREC_F64
    uf, vf, wf, _ = O.multiscale(im1, d1, im2, d2, sigma, hpb, po)
    e = np.sqrt((ub - uf) ** 2 + (vb - vf) ** 2)
    print(f"{name} seed {seed}: |GPU-ref| u {du:.3g} v {dv:.3g} w {dw:.3g}; |GPU-f64| flow mean {e.mean():.3g} "
          f"median {np.median(e):.3g} p99 {np.percentile(e, 99):.3g} max {e.max():.3g} px, "
          f"w max {_maxdiff(wb, wf):.3g} m")
    return (ua, va, wa), (ub, vb, wb), pair


def test_multiscale_320x240(ctx, R):
    _pair_parity(ctx, R, "320x240_10x30", seed=11)


def test_multiscale_320x240_no_exits(ctx, R):
    _pair_parity(ctx, R, "320x240_10x30", seed=12, exits=False)


def test_multiscale_640x480(ctx, R):
    ref, got, pair = _pair_parity(ctx, R, "640x480_15x50", seed=21)
    gt = pair["flow"]
    epe = np.sqrt((got[0] - gt[..., 0]) ** 2 + (got[1] - gt[..., 1]) ** 2)
    # the solve recovers the synthetic motion (occluded slivers at segment
    # boundaries have no correspondence and drift, so the check is robust)
    assert np.median(epe) < 0.2, np.median(epe)


def test_sigma_on_device_matches_given(ctx, R):
    """pair_load(sigma1=NULL) computes the stdevs on the device."""
    O = oracle()
    W, H = 160, 120
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 13, 4)
    hpb, lv = hpb_for(pair, W, H, 0.5, 20)
    pg = R.make_params(W, H, n_outer=4, n_sor=25, min_width=20)
    a = ctx.multiscale(pg, im1, d1, im2, d2, sigma, hpb)
    b = ctx.multiscale(pg, im1, d1, im2, d2, None, hpb)
    for x, y in zip(a[:3], b[:3]):
        assert _maxdiff(x, y) <= 1e-6


def test_with_init(ctx, R):
    O = oracle()
    W, H = 128, 96
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 14, 4)
    hpb, lv = hpb_for(pair, W, H, 0.5, 16)
    hpb0 = hpb[: W * H * 2]
    est = pair["flow"].astype(np.float64) * 0.9
    po = O.make_params(W, H, n_outer=3, n_sor=40)
    pg = R.make_params(W, H, n_outer=3, n_sor=40)
    ua, va, wa, sa = O.with_init(im1, d1, im2, d2, sigma, hpb0, est, po)
    ub, vb, wb, sb = ctx.with_init(pg, im1, d1, im2, d2, est, sigma, hpb0)
    assert sb.sweeps_exec[0] == sa.sweeps_exec[0]
    assert _maxdiff(ua, ub) <= TOL_UV and _maxdiff(va, vb) <= TOL_UV and _maxdiff(wa, wb) <= TOL_W


def test_regularisation_types(ctx, R):
    O = oracle()
    W, H = 128, 96
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 15, 4)
    for reg in (O.REG_ISOTROPIC, O.REG_LETOUZEY):
        po = O.make_params(W, H, n_outer=3, n_sor=25, min_width=32, reg=reg)
        pg = R.make_params(W, H, n_outer=3, n_sor=25, min_width=32, reg=reg)
        ua, va, wa, _ = O.multiscale(im1, d1, im2, d2, sigma, None, po)
        ub, vb, wb, _ = ctx.multiscale(pg, im1, d1, im2, d2, sigma, None)
        assert _maxdiff(ua, ub) <= TOL_UV and _maxdiff(va, vb) <= TOL_UV and _maxdiff(wa, wb) <= TOL_W


def test_error_paths(R):
    with R.Context(0, hbm_budget=1 << 20) as c:      # 1 MiB budget
        W, H = 320, 240
        p = R.make_params(W, H, reg=R.REG_ISOTROPIC)
        z = np.zeros((H, W, 3))
        d = np.ones((H, W))
        with pytest.raises(R.RgbdFlowError) as e:
            c.pair_load(p, z, d, z, d)
        assert e.value.code == R.ERR_BUDGET
    with R.Context(0) as c:
        # HPB without maps: made on the device, which needs the depth-map and
        # colour pyramids to have equal level sizes (47 x 33 at ratio 0.7: 11 x 7
        # against 11 x 8 at the coarsest level)
        p = R.make_params(47, 33, ratio=0.7, min_width=11)
        z = np.zeros((33, 47, 3))
        d = np.ones((33, 47))
        with pytest.raises(R.RgbdFlowError) as e:
            c.pair_load(p, z, d, z, d)
        assert e.value.code == R.ERR_UNSUPPORTED
        z = np.zeros((48, 64, 3))
        d = np.ones((48, 64))
        p = R.make_params(64, 48, reg=R.REG_ISOTROPIC, n_outer=5000)
        c.pair_load(p, z, d, z, d)
        with pytest.raises(R.RgbdFlowError) as e:
            c.pair_run(p)
        assert e.value.code == R.ERR_ARG


@pytest.mark.parametrize("n_inner", [0, 2, 3])
@pytest.mark.parametrize("variant", ["plain", "tiled"])
def test_multiscale_inner_iterations(ctx, R, n_inner, variant):
    """nInner other than 1: the first inner iteration's fused pass stores the
    derivatives for the later ones (and nInner = 0 only re-warps)."""
    O = oracle()
    W, H = 128, 96
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 16, 4)
    hpb, lv = hpb_for(pair, W, H, 0.5, 32)
    kw = dict(n_outer=3, n_inner=n_inner, n_sor=30, min_width=32)
    ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb, O.make_params(W, H, **kw))
    pg = R.make_params(W, H, relax_variant=R.RELAX_TILED if variant == "tiled" else R.RELAX_PLAIN, **kw)
    ub, vb, wb, sb = ctx.multiscale(pg, im1, d1, im2, d2, sigma, hpb)
    L = sa.n_levels
    assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)
