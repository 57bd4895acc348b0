"""Pins of the oracle (the CPU restatement used as the checker).

The reference ships no tests, golden vectors or fixtures, and cannot be built
here (CUDA/thrust, Boost, OpenCV, PCL), so each oracle function is pinned by
what the mathematics fixes: closed forms, invariants, special cases that
reduce to a library routine (scipy), and direct solves.  Each check is chosen
so that a plausible mistake (wrong tap order or index, a dropped term, a
transposed neighbour, a sign) fails it.
"""
import math

import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import level_inputs, make_case, oracle

O = oracle()


# ------------------------------------------------------------------ pyramid
def test_pyramid_level_sizes():
    assert O.pyramid_levels(640, 480, 0.5, 40) == [(640, 480), (320, 240), (160, 120), (80, 60), (40, 30)]
    assert O.pyramid_levels(320, 240, 0.5, 40) == [(320, 240), (160, 120), (80, 60), (40, 30)]
    assert O.pyramid_levels(960, 540, 0.5, 60) == [(960, 540), (480, 270), (240, 135), (120, 67), (60, 33)]
    assert O.pyramid_levels(1280, 720, 0.5, 80) == [(1280, 720), (640, 360), (320, 180), (160, 90), (80, 45)]
    # the reference's demo value minWidth = 60 gives 4 levels at 640 (runFlowTwoFrames.cpp:72)
    assert len(O.pyramid_levels(640, 480, 0.5, 60)) == 4
    # ratios outside [0.4, 0.98] fall back to 0.75
    assert O.pyramid_levels(100, 80, 0.2, 30) == O.pyramid_levels(100, 80, 0.75, 30)


def _taps(sigma, fs):
    g = [math.exp(-(i * i) / (sigma * sigma * 2)) for i in range(-fs, fs + 1)]
    s = 0.0
    for x in g:
        s += x
    return np.array([x / s for x in g])


@pytest.mark.parametrize("sigma,fs", [(1.0, 3), (2.0, 6), (0.7, 2)])
def test_gaussian_impulse_response(sigma, fs):
    """Separable filtering of an impulse returns the outer product of the
    normalised taps exp(-i^2 / 2 sigma^2) (exact: one product per sample)."""
    img = np.zeros((31, 33))
    img[15, 16] = 1.0
    out = O.gaussian_smooth(img, sigma, fs)
    g = _taps(sigma, fs)
    exp = np.zeros_like(img)
    exp[15 - fs:15 + fs + 1, 16 - fs:16 + fs + 1] = np.outer(g, g)
    assert np.array_equal(out, exp)


def test_gaussian_constant_and_ramp():
    c = np.full((40, 50, 3), 0.37)
    out = O.gaussian_smooth(c, 2.0, 6)
    assert np.max(np.abs(out - 0.37)) < 1e-15
    yy, xx = np.mgrid[0:40, 0:50].astype(float)
    ramp = 0.01 * xx + 0.02 * yy + 0.5
    out = O.gaussian_smooth(ramp, 2.0, 6)
    inner = (slice(6, -6), slice(6, -6))    # symmetric taps keep a linear function
    assert np.max(np.abs(out[inner] - ramp[inner])) < 1e-13
    assert not np.allclose(out[0, 0], ramp[0, 0])  # clamp-to-edge bends it at the border


def test_resize_mapping():
    """ResizeImage reads (j+1)/r - 1: ratio 0.5 picks odd pixels exactly."""
    yy, xx = np.mgrid[0:24, 0:32].astype(float)
    ramp = 3.0 * xx + 100.0 * yy
    half = O.resize_ratio(ramp, 0.5)
    assert half.shape == (12, 16)
    assert np.array_equal(half, ramp[1::2, 1::2])
    quarter = O.resize_ratio(ramp, 0.25)
    assert np.array_equal(quarter, ramp[3::4, 3::4])
    # upsample by dims: x = (j+1)*Ws/Wd - 1 = (j-1)/2, truncation toward zero, clamp
    up = O.resize_dims(ramp[1::2, 1::2], 32, 24)
    src = ramp[1::2, 1::2]
    assert up[0, 0] == src[0, 0] and up[1, 1] == src[0, 0]
    assert np.isclose(up[2, 2], 0.25 * (src[0, 0] + src[0, 1] + src[1, 0] + src[1, 1]))
    x = (np.arange(32) - 1) / 2.0
    exp_row = 3.0 * (2 * np.clip(x, 0, 15) + 1)  # linear in the source grid
    assert np.allclose(up[1, :31] - 100.0, exp_row[:31])


def test_pyramid_of_constant_and_levels():
    img = np.full((60, 80), 1.25)
    P = O.gaussian_pyramid(img, 0.5, 12)
    assert [p.shape for p in P] == [(60, 80), (30, 40), (15, 20)]
    for p in P:
        assert np.max(np.abs(p - 1.25)) < 1e-14
    rng = np.random.default_rng(0)
    img = rng.uniform(size=(64, 96))
    P = O.gaussian_pyramid(img, 0.5, 10)
    assert np.array_equal(P[0], img)                                   # level 0 unsmoothed
    assert np.array_equal(P[1], O.resize_ratio(O.gaussian_smooth(img, 1.0, 3), 0.5))
    assert np.array_equal(P[2], O.resize_ratio(O.gaussian_smooth(img, 2.0, 6), 0.25))
    assert np.array_equal(P[3], O.resize_ratio(O.gaussian_smooth(P[1], 2.0, 6), 0.25))  # re-smooth level i-2


# ----------------------------------------------------------------- features
def test_im2feature_closed_forms():
    H, W = 20, 30
    xx = np.mgrid[0:H, 0:W][1].astype(float)
    b, g, r = 0.1 + 0.001 * xx ** 4 / 1e3, 0.2 + 0.0 * xx, 0.3 + 0.002 * xx ** 3 / 1e2
    bgr = np.stack([b, g, r], -1)
    f = O.im2feature(bgr)
    gray = 0.114 * b + 0.587 * g + 0.299 * r
    assert np.allclose(f[..., 0], gray, rtol=0, atol=1e-15)
    # the 5-tap [1,-8,0,8,-1]/12 derivative is exact on quartics away from the border
    dgray = 0.114 * 0.004 * xx ** 3 / 1e3 + 0.299 * 0.006 * xx ** 2 / 1e2
    assert np.allclose(f[:, 2:-2, 1], dgray[:, 2:-2], rtol=0, atol=1e-12)
    assert np.allclose(f[..., 2], 0, atol=1e-15)       # no y variation
    assert np.array_equal(f[..., 3], g - b) and np.array_equal(f[..., 4], g - r)


# -------------------------------------------------------------------- warps
def test_warp_cpu_identities():
    rng = np.random.default_rng(1)
    H, W = 17, 23
    I1, I2 = rng.uniform(size=(H, W, 5)), rng.uniform(size=(H, W, 5))
    D1, D2 = rng.uniform(1, 3, (H, W)), rng.uniform(1, 3, (H, W))
    z = np.zeros((H, W))
    wI, wD = O.warp_cpu(I1, D1, I2, D2, z, z, z)
    assert np.array_equal(wI, I2) and np.array_equal(wD, D2)
    one = np.ones((H, W))
    wI, wD = O.warp_cpu(I1, D1, I2, D2, one, z, 0.25 * one)
    assert np.array_equal(wI[:, :-1], I2[:, 1:])                     # integer shift
    assert np.array_equal(wI[:, -1], I1[:, -1])                      # out of bounds -> frame 1
    assert np.array_equal(wD[:, :-1], D2[:, 1:] - 0.25)
    assert np.array_equal(wD[:, -1], D1[:, -1] - 0.25)               # ... minus w as well
    # bilinear interpolation reproduces a linear image
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    lin = 0.3 * xx - 0.2 * yy + 1
    u, v = rng.uniform(-2, 2, (H, W)), rng.uniform(-2, 2, (H, W))
    I = np.repeat(lin[..., None], 5, -1)
    wI, wD = O.warp_cpu(I, lin, I, lin, u, v, z)
    x, y = xx + u, yy + v
    inb = (x >= 0) & (x <= W - 1) & (y >= 0) & (y <= H - 1)
    assert np.allclose(wD[inb], (0.3 * x - 0.2 * y + 1)[inb], atol=1e-13)
    assert np.array_equal(wD[~inb], lin[~inb])


# ------------------------------------------------------------------ weights
def test_stereo_error_and_delta():
    se1 = 1 / (1 / .00285 + 1)
    assert abs(se1 - 0.0028419) < 1e-7
    d = np.full((12, 14), 1.5, np.float32)
    sig = O.depth_stdevs(d, 4)
    assert np.allclose(sig, 1.5 ** 2 / (1 / .00285 + 1.5), rtol=1e-6)   # 0.0063852
    assert abs(float(sig[0, 0]) - 0.0063852) < 1e-7
    delta = O.depth_data_weights(sig)
    assert np.allclose(delta, 0.4451, atol=1e-4)


def test_l1_distance_matches_scipy():
    pair, im1, d1, im2, d2, sigma = make_case(64, 48, 2, 4)
    d = d1.astype(np.float32)
    d[5:9, 10:20] = 0
    dist = O.l1_distance_to_boundary(d, 4)
    # boundary mask, written from staticDepthNoiseModeling.cpp:95-97 independently
    se = lambda z: z * z / (1 / .00285 + z)
    ratio = lambda z: se(np.float64(z)) / se(1.0)
    H, W = d.shape
    mask = np.zeros((H, W), bool)
    dn = d[1:, :]
    mask[:-1, :] |= (dn > 0) & (np.abs(d[:-1, :] - dn) > .02 * ratio(np.minimum(d[:-1, :], dn)))
    dr = d[:, 1:]
    mask[:, :-1] |= (dr > 0) & (np.abs(d[:, :-1] - dr) > .02 * ratio(np.minimum(d[:, :-1], dr)))
    ref = ndi.distance_transform_cdt(~mask, metric="taxicab").astype(np.float32)
    ref = np.minimum(ref, 5)
    assert np.array_equal(dist, ref)


def test_sigma_inflation_profile():
    H, W = 20, 40
    d = np.full((H, W), 1.0, np.float32)
    d[:, 20:] = 2.0                                   # a vertical depth edge between columns 19|20
    sig = O.depth_stdevs(d, 4)
    base1 = 1 / (1 / .00285 + 1)
    assert np.isclose(sig[10, 19], 10 * base1, rtol=1e-6)              # on the boundary: x10
    assert np.isclose(sig[10, 17], base1 * 10 ** 0.5, rtol=1e-6)       # d = 2: x10^(1/2)
    assert np.isclose(sig[10, 15], base1, rtol=1e-6)                   # d = 4: unchanged
    assert np.isclose(sig[10, 5], base1, rtol=1e-6)
    invalid = np.zeros((H, W), np.float32)
    assert np.all(O.depth_stdevs(invalid, 4) == 1.0)                   # all invalid: sigma 1


def test_depth_magnitude_weights_quirk():
    f = 525.0
    D = np.full((30, 40), 2.0)
    D[0, 0] = 0.0
    m = O.depth_mag_weights(D, 640, 480, f)          # level 40x30 under the full-size camera
    idx = (np.arange(1200) // 640) * 40 + np.arange(1200) % 640
    src = D.ravel()[idx]
    exp = np.where(src > 0, (f / np.where(src > 0, src, 1)) ** 2, (np.float32(f) / 2) ** 2)
    assert np.array_equal(m.ravel(), exp)
    full = O.depth_mag_weights(D, 40, 30, f)          # at full resolution it is plain (f/D)^2
    assert full[0, 0] == (f / 2) ** 2 and np.all(full.ravel()[1:] == (f / 2.0) ** 2)


def test_regularisation_weights():
    rng = np.random.default_rng(3)
    H, W = 6, 7
    hpb = rng.uniform(size=(H, W, 2)).astype(np.float32)
    D = rng.uniform(1, 1.1, (H, W))
    r = O.reg_weights(hpb, D, O.REG_HPB)
    assert np.array_equal(r[:, :-1, 0], 1 - hpb[:, :-1, 0].astype(np.float64) * .99)
    assert np.all(r[:, -1, 0] == 0) and np.all(r[-1, :, 1] == 0)
    assert np.all(O.reg_weights(None, D, O.REG_ISOTROPIC)[:, :-1, 0] == 1)
    let = O.reg_weights(None, D, O.REG_LETOUZEY)
    assert np.allclose(let[:-1, :, 1], np.exp(-((D[:-1] - D[1:]) / .02) ** 2))


def test_median_fill():
    d = np.full((8, 9), 2.0)
    assert np.array_equal(O.median_fill(d), d)                        # no holes: no-op
    d[3, 4] = 0
    assert np.all(O.median_fill(d) == 2.0)
    g = np.arange(9, dtype=float).reshape(3, 3) + 1.1
    g[1, 1] = 0
    out = O.median_fill(g)
    nb = np.sort(np.delete(g.ravel(), 4).astype(np.float32))
    assert out[1, 1] == nb[4]                                         # upper median, via f32


def test_depth_map_pyramid():
    d = np.full((32, 48), 1.7)
    P = O.depth_map_pyramid(d, 0.5, 10)
    assert all(np.allclose(p, 1.7, atol=1e-6) for p in P)
    d[:8, :8] = -1
    P = O.depth_map_pyramid(d, 0.5, 10)
    assert np.all(P[1][:4, :4] == -1) and np.all(P[1][5:, 5:] > 0)
    e = np.full((32, 48), 1.0)
    e[:, 25:] = 2.0                                  # an edge inside a 2x2 block is not averaged across
    P = O.depth_map_pyramid(e, 0.5, 10)
    assert set(np.unique(np.round(P[1], 6))) <= {1.0, 2.0}


# --------------------------------------------------------- system and solve
def _sparse_system(coef):
    """M x = b of the relaxation's fixed point, assembled from the compact
    coefficients: row (p,k) has B_kk' on the pixel's own unknowns and the
    edge weight of each existing neighbour on the neighbour's unknown k."""
    _, H, W = coef.shape
    N = H * W
    c = coef.reshape(13, N)
    eRxy, eRz, eDxy, eDz, Buu, Bvv, Bww, Buv, Buw, Bvw = c[:10]
    rows, cols, vals = [], [], []
    B = [[Buu, Buv, Buw], [Buv, Bvv, Bvw], [Buw, Bvw, Bww]]
    idx = lambda p, k: k * N + p
    for p in range(N):
        i, j = divmod(p, W)
        for k in range(3):
            for kk in range(3):
                rows.append(idx(p, k)); cols.append(idx(p, kk)); vals.append(B[k][kk][p])
            eR, eD = (eRxy, eDxy) if k < 2 else (eRz, eDz)
            for (ok, q, e) in ((j > 0, p - 1, eR[p - 1] if j > 0 else 0), (i > 0, p - W, eD[p - W] if i > 0 else 0),
                               (j < W - 1, p + 1, eR[p]), (i < H - 1, p + W, eD[p])):
                if ok:
                    rows.append(idx(p, k)); cols.append(idx(q, k)); vals.append(e)
    M = sp.csr_matrix((vals, (rows, cols)), shape=(3 * N, 3 * N))
    b = np.concatenate([c[10], c[11], c[12]])
    return M, b


def test_system_is_symmetric_negative_definite():
    li, (u, v, w), _ = level_inputs(seed=4, W=14, H=12)
    coef = O.level_system(li, O.make_params(14, 12, precision=O.PREC_F64), u, v, w)
    M, b = _sparse_system(coef)
    Md = M.toarray()
    assert np.allclose(Md, Md.T, rtol=1e-12, atol=0)
    assert np.max(np.linalg.eigvalsh(Md)) < 0           # -M is SPD


def test_relaxation_converges_to_direct_solve():
    """At convergence (clamp inactive) every ordering reaches M x = b; a
    transposed neighbour, a sign error or a missing coupling would not."""
    li, (u, v, w), _ = level_inputs(seed=5, W=14, H=12)
    coef = O.level_system(li, O.make_params(14, 12, precision=O.PREC_F64), u, v, w)
    coef[10:13] *= 1e-2                                  # keep every step inside the 0.3 clamp
    M, b = _sparse_system(coef)
    x_direct = spla.spsolve(M.tocsc(), b)
    x = O.relax(coef, 4000, np.zeros((3, 12, 14)), precision=O.PREC_F64)
    assert np.max(np.abs(x.reshape(-1) - x_direct)) < 1e-9 * max(1.0, np.abs(x_direct).max())


def test_relaxation_step_clamp():
    li, (u, v, w), _ = level_inputs(seed=6, W=16, H=12)
    coef = O.level_system(li, O.make_params(16, 12), u, v, w)
    coef[10:13] = 1e4                                    # a huge rhs: every update is clamped
    x = O.relax(coef, 1, np.zeros((3, 12, 16)))
    assert np.max(np.abs(x)) <= np.float32(0.3) * (1 + 1e-6)
    assert np.isclose(np.max(np.abs(x)), 0.3, rtol=1e-6)


def test_identical_frames_give_zero_flow():
    pair, im1, d1, im2, d2, sigma = make_case(96, 72, 7, 3)
    p = O.make_params(96, 72, reg=O.REG_ISOTROPIC, n_outer=3, n_sor=25, min_width=24)
    u, v, w, st = O.multiscale(im1, d1, im1, d1, sigma, None, p)
    assert np.all(u == 0) and np.all(v == 0) and np.all(w == 0)


def test_multiscale_recovers_synthetic_motion():
    pair, im1, d1, im2, d2, sigma = make_case(160, 120, 8, 4)
    from helpers import hpb_for
    hpb, lv = hpb_for(pair, 160, 120, 0.5, 20)
    p = O.make_params(160, 120, min_width=20, n_outer=10, n_sor=30)
    u, v, w, st = O.multiscale(im1, d1, im2, d2, sigma, hpb, p)
    gt = pair["flow"]
    epe = np.sqrt((u - gt[..., 0]) ** 2 + (v - gt[..., 1]) ** 2).mean()
    zero = np.sqrt(gt[..., 0] ** 2 + gt[..., 1] ** 2).mean()
    assert epe < 0.5 * zero and epe < 1.0


def test_reference_vs_f64_precision_close():
    """At equal iteration counts the two precisions differ by f32 rounding only.
    (With the exits on they may not run equal counts: a step clamped to +-0.3
    changes x by 0.3 rounded in the working precision, and the divergence exit
    `maxdiff >= 0.3` (.cu:921) fires on that last bit in f32 but not in f64.)"""
    li, (u, v, w), _ = level_inputs(seed=9, W=48, H=36)
    kw = dict(n_outer=3, n_sor=50, disable_inner_exit=True, disable_outer_exit=True)
    a = O.level_solve(li, O.make_params(48, 36, **kw), True, u, v, w)
    b = O.level_solve(li, O.make_params(48, 36, precision=O.PREC_F64, **kw), True, u, v, w)
    assert np.max(np.abs(a[0] - b[0])) < 1e-4 and np.max(np.abs(a[2] - b[2])) < 1e-5
    c = O.level_solve(li, O.make_params(48, 36, n_outer=2, n_sor=25), True, u, v, w)
    d = O.level_solve(li, O.make_params(48, 36, n_outer=2, n_sor=25, precision=O.PREC_F64), True, u, v, w)
    assert c[3].last_maxdiff >= np.float32(0.3) > d[3].last_maxdiff - 1e-7


# ------------------------------------- the level-solve warp and the exits
def test_warp_gpu_style_identities():
    rng = np.random.default_rng(11)
    H, W = 13, 17
    I1, I2 = rng.uniform(size=(H, W, 5)), rng.uniform(size=(H, W, 5))
    D1, D2 = rng.uniform(1, 3, (H, W)), rng.uniform(1, 3, (H, W))
    z = np.zeros((H, W))
    for prec in (O.PREC_REF, O.PREC_F64):
        wI, wD = O.warp_gpu_style(I1, D1, I2, D2, z, z, z, prec)
        assert np.array_equal(wI, I2) and np.array_equal(wD, D2)
        wI, wD = O.warp_gpu_style(I1, D1, I2, D2, -np.ones((H, W)), z, 0.5 * np.ones((H, W)), prec)
        assert np.array_equal(wI[:, 1:], I2[:, :-1]) and np.array_equal(wI[:, 0], I1[:, 0])
        assert np.array_equal(wD[:, 1:], D2[:, :-1] - 0.5) and np.array_equal(wD[:, 0], D1[:, 0] - 0.5)
    # f32 bilinear weights (REF) vs f64 weights: both reproduce a linear image, to f32 accuracy
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    lin = 0.3 * xx - 0.2 * yy + 1
    u, v = rng.uniform(-2, 2, (H, W)), rng.uniform(-2, 2, (H, W))
    x, y = xx + u, yy + v
    inb = (x >= 0) & (x <= W - 1) & (y >= 0) & (y <= H - 1)
    I = np.repeat(lin[..., None], 5, -1)
    _, a = O.warp_gpu_style(I, lin, I, lin, u, v, z, O.PREC_F64)
    _, b = O.warp_gpu_style(I, lin, I, lin, u, v, z, O.PREC_REF)
    exact = 0.3 * x - 0.2 * y + 1
    assert np.allclose(a[inb], exact[inb], rtol=0, atol=1e-12)
    assert np.allclose(b[inb], exact[inb], rtol=0, atol=1e-5)
    assert not np.allclose(b[inb], exact[inb], rtol=0, atol=1e-12)   # the weights really are f32


def test_exits_on_a_converged_system():
    """Identical frames give a zero system: the relaxation stops at the first
    check (iteration 25, max change 0 < 1e-5) and, where allowed, the outer
    loop after its first iteration (statistic 0 < 4e-4)."""
    li, (u, v, w), _ = level_inputs(seed=13, W=32, H=24, level_flow=False)
    li2 = O.LevelInputs(li.I1, li.I1, li.I1, li.D1, li.D1, li.D1, li.delta, li.s, li.m, li.r, li.focal_level)
    z = np.zeros((24, 32))
    _, _, _, st = O.level_solve(li2, O.make_params(32, 24, n_outer=5, n_sor=60), True, z, z, z)
    assert (st.outer_exec, st.sweeps_exec) == (1, 25)
    _, _, _, st = O.level_solve(li2, O.make_params(32, 24, n_outer=5, n_sor=60), False, z, z, z)
    assert (st.outer_exec, st.sweeps_exec) == (5, 125)
    _, _, _, st = O.level_solve(li2, O.make_params(32, 24, n_outer=5, n_sor=60, disable_inner_exit=True,
                                                    disable_outer_exit=True), True, z, z, z)
    assert (st.outer_exec, st.sweeps_exec) == (5, 300)


# ------------------------------------------- assembly vs the lagged energy
def _derivs(A, B):
    """x/y derivatives of 0.4 A + 0.6 B with [1,-8,0,8,-1]/12, clamp-to-edge."""
    M = 0.4 * A + 0.6 * B
    H, W = M.shape[:2]
    g = np.array([1, -8, 0, 8, -1]) / 12.0
    dx = sum(g[m] * M[:, np.clip(np.arange(W) - 2 + m, 0, W - 1)] for m in range(5))
    dy = sum(g[m] * M[np.clip(np.arange(H) - 2 + m, 0, H - 1)] for m in range(5))
    return dx, dy


def test_system_is_the_lagged_energy_gradient():
    """With q = 1 (no occlusion), isotropic regularisation and a constant
    depth-smoothness weight, the reference's quirks (q^2 on the It products
    only, the Laplacian normalisation, edge-owner weights) vanish and the
    assembled system is exactly the linearised Euler-Lagrange system of the
    quadratic energy with frozen robust weights:
        b - M du = grad E(du),
        E(du) = 1/2 sum_p [ <psi (It + Ix du + Iy dv)^2> + pi (Dt + Dx du + Dy dv - dw)^2
                            + gamma ((u+du)^2 + (v+dv)^2 + m (w+dw)^2) ]
              + alpha/2 sum_edges phi_o r_o [ d(u+du)^2 + d(v+dv)^2 + s d(w+dw)^2 ].
    A dropped term, a wrong sign or a transposed neighbour breaks the equality."""
    rng = np.random.default_rng(14)
    H, W = 6, 8
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    I1 = np.stack([np.sin(0.5 * xx + k) * np.cos(0.3 * yy - k) + 0.1 * k for k in range(5)], -1)
    I2 = I1 + 0.05 * rng.normal(size=I1.shape)
    Iw = I1 + 0.05 * rng.normal(size=I1.shape)
    D1 = 1.5 + 0.002 * xx + 0.003 * yy
    D2 = D1 + 1.0                                     # q = 1 everywhere
    Dw = D1 + 0.002 * rng.normal(size=D1.shape)
    s_const, alpha, gamma = 3.0, 0.7, 0.05
    delta = rng.uniform(0.5, 1.5, (H, W))
    m = rng.uniform(1, 2, (H, W))
    r = O.reg_weights(None, D1, O.REG_ISOTROPIC)
    u, v, w = (0.1 * rng.normal(size=(H, W)) for _ in range(3))
    li = O.LevelInputs(I1, I2, Iw, D1, D2, Dw, delta, np.full((H, W), s_const), m, r, np.float32(1.0))
    p = O.make_params(W, H, alpha=alpha, gamma=gamma, precision=O.PREC_F64, reg=O.REG_ISOTROPIC)
    coef = O.level_system(li, p, u, v, w)
    M, b = _sparse_system(coef)
    phi, psi, pi = O.level_weights(li, p, u, v, w)
    Ix, Iy = _derivs(I1, Iw)
    It = Iw - I1
    Dx, Dy = _derivs(D1, Dw)
    Dt = Dw - D1
    assert max(np.abs(Dx).max(), np.abs(Dy).max(), np.abs(Dt).max()) < 0.04   # clamp inactive
    N = H * W

    def energy(x):
        du, dv, dw = x[:N].reshape(H, W), x[N:2 * N].reshape(H, W), x[2 * N:].reshape(H, W)
        rc = It + Ix * du[..., None] + Iy * dv[..., None]
        e = 0.5 * np.sum(np.mean(psi * rc ** 2, -1))
        e += 0.5 * np.sum(pi * (Dt + Dx * du + Dy * dv - dw) ** 2)
        e += 0.5 * gamma * np.sum((u + du) ** 2 + (v + dv) ** 2 + m * (w + dw) ** 2)
        U, V, Wz = u + du, v + dv, w + dw
        wr = (phi * r[..., 0])[:, :-1]
        wd = (phi * r[..., 1])[:-1, :]
        for F, sc in ((U, 1.0), (V, 1.0), (Wz, s_const)):
            e += 0.5 * alpha * sc * (np.sum(wr * np.diff(F, axis=1) ** 2) + np.sum(wd * np.diff(F, axis=0) ** 2))
        return e

    for trial in range(3):
        x = 0.05 * rng.normal(size=3 * N)
        grad = np.empty(3 * N)
        for i in range(3 * N):
            e = np.zeros(3 * N)
            e[i] = 1.0
            grad[i] = (energy(x + e) - energy(x - e)) / 2.0
        lhs = b - M @ x
        assert np.max(np.abs(lhs - grad)) < 1e-9 * max(1.0, np.abs(grad).max()), np.max(np.abs(lhs - grad))


# ------------------------------------------------- FP64 residual and energy
def test_system_residual_matches_sparse_operator():
    """The residual test's f64 reduction equals ||b - M x|| / ||b|| and the
    energy b'x - x'Mx/2 with M assembled independently (scipy); a transposed
    neighbour, a dropped coupling or a sign error changes both."""
    li, (u, v, w), _ = level_inputs(seed=21, W=18, H=14)
    coef = O.level_system(li, O.make_params(18, 14), u, v, w)
    M, b = _sparse_system(coef)
    x = np.random.default_rng(3).normal(scale=0.05, size=(3, 14, 18)).astype(np.float32).astype(np.float64)
    rel, en = O.system_residual(coef, x)
    xv = x.reshape(-1)
    r = b - M @ xv
    assert np.isclose(rel, np.linalg.norm(r) / np.linalg.norm(b), rtol=1e-12, atol=0)
    assert np.isclose(en, b @ xv - 0.5 * xv @ (M @ xv), rtol=1e-12, atol=0)
    # at the direct solution the residual vanishes; the same x against the
    # system with its right and down couplings swapped (a planted indexing
    # mistake) or with the rhs negated (a sign error) does not
    xs = spla.spsolve(M.tocsc(), b).reshape(3, 14, 18)
    assert O.system_residual(coef, xs)[0] < 1e-10
    bad = coef.copy()
    bad[0], bad[2] = coef[2], coef[0]
    assert O.system_residual(bad, xs)[0] > 1e-3
    bad = coef.copy()
    bad[10:13] *= -1
    assert O.system_residual(bad, xs)[0] > 1.0
    # zero increment: residual 1, energy 0
    assert O.system_residual(coef, np.zeros_like(x)) == (1.0, 0.0)


def test_relax_residual_reaches_the_minimiser_for_any_factor():
    """Relaxing until the relative residual is <= tol reaches the energy
    minimum of the direct solve.  In the point SOR order (u, then v from the
    new u, then w) every factor in (0, 2) converges, so omega = 0.7 and
    omega = 1.8 agree on the converged energy; the test runs only at
    iterations 25, 50, ... and stops at the first that passes."""
    li, (u, v, w), _ = level_inputs(seed=22, W=20, H=16)
    coef = O.level_system(li, O.make_params(20, 16, precision=O.PREC_F64), u, v, w)
    coef[10:13] *= 1e-2                                  # every step inside the 0.3 clamp
    M, b = _sparse_system(coef)
    x_direct = spla.spsolve(M.tocsc(), b)
    e_min = b @ x_direct - 0.5 * x_direct @ (M @ x_direct)
    x0 = np.zeros((3, 16, 20))
    energies = {}
    for gs, om in ((False, 0.7), (True, 0.7), (True, 1.8)):
        x, it, res, en = O.relax_residual(coef, 20000, 1e-8, x0, omega=om, precision=O.PREC_F64, pixel_gs=gs)
        assert it % 25 == 0 and res <= 1e-8
        _, _, res_prev, _ = O.relax_residual(coef, it - 25, 0.0, x0, omega=om, precision=O.PREC_F64, pixel_gs=gs)
        assert res_prev > 1e-8                           # the previous check had not passed
        assert abs(en - e_min) <= 1e-6 * abs(e_min)
        assert np.max(np.abs(x.reshape(-1) - x_direct)) < 1e-6 * np.abs(x_direct).max()
        energies[(gs, om)] = (it, en)
    assert energies[(True, 1.8)][0] < energies[(True, 0.7)][0] / 3   # over-relaxation pays
    assert abs(energies[(True, 1.8)][1] - energies[(False, 0.7)][1]) <= 1e-4 * abs(e_min)
    # the reference's order (each unknown from the others' old values) does
    # not converge with omega = 1.8 on this system: it stalls far away
    _, _, res_j, _ = O.relax_residual(coef, 2000, 0.0, x0, omega=1.8, precision=O.PREC_F64, pixel_gs=False)
    assert res_j > 1.0


def test_level_solve_residual_exit():
    """With the residual exit each relaxation stops at the first check whose
    residual is <= tol; the stats report it.  A tolerance of 0 never stops it."""
    li, (u, v, w), _ = level_inputs(seed=23, W=32, H=24)
    kw = dict(n_outer=3, n_sor=200)
    _, _, _, st = O.level_solve(li, O.make_params(32, 24, residual_exit=True, residual_tol=1e-3, omega=1.8,
                                                  pixel_gs=True, **kw), False, u, v, w)
    assert st.residual_checks >= 3 and 0 <= st.last_residual <= 1e-3
    assert st.sweeps_exec % 25 == 0 and st.sweeps_exec < 3 * 200
    _, _, _, st0 = O.level_solve(li, O.make_params(32, 24, residual_exit=True, residual_tol=0.0, **kw), False,
                                 u, v, w)
    assert st0.sweeps_exec == 3 * 200 and st0.residual_checks == 3 * 8 and st0.last_residual > 0


def _img_values(rng):
    """Feature-like values plus the hard cases of a 16-bit rounding: FP16
    subnormals, exact ties between neighbouring FP16 / BF16 numbers, the FP16
    overflow boundary, zeros and both signs."""
    v = [rng.uniform(-1, 1, 2000), rng.uniform(-1e-3, 1e-3, 500), rng.uniform(-1e-6, 1e-6, 500),
         np.ldexp(rng.integers(1, 2048, 300).astype(np.float64) + 0.5, -24),     # FP16 subnormal ties
         np.ldexp(rng.integers(1024, 2048, 300).astype(np.float64) + 0.5, -10),  # FP16 ties in [1, 2)
         np.ldexp(rng.integers(128, 256, 300).astype(np.float64) + 0.5, -7),     # BF16 ties in [1, 2)
         np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, -65520.0, 1e5, 2.0 ** -25, 2.0 ** -24, 3 * 2.0 ** -26])]
    x = np.concatenate(v)
    return np.concatenate([x, -x])


def test_image_storage_rounding():
    """The feature-image storage variants round like numpy's float16 and
    torch's bfloat16 conversions of the f32 value (round to nearest even,
    FP16 subnormals and overflow included); f64 storage leaves values as they
    are."""
    import torch
    O = oracle()
    x = _img_values(np.random.default_rng(5))
    f32 = x.astype(np.float32)
    want16 = f32.astype(np.float16).astype(np.float64)
    got16 = O.img_round(x, O.IMG_F16)
    assert np.array_equal(got16, want16, equal_nan=True)
    wantb = torch.from_numpy(f32).to(torch.bfloat16).to(torch.float64).numpy()
    gotb = O.img_round(x, O.IMG_BF16)
    assert np.array_equal(gotb, wantb)
    assert np.array_equal(O.img_round(x, O.IMG_F64), x)
    # a planted mistake fails: truncation instead of rounding differs on the ties
    assert not np.array_equal(np.trunc(f32.astype(np.float64) * 2.0 ** 10) / 2.0 ** 10, got16)


def _plane(W, H, rng):
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    return 1.5 + 0.002 * xs - 0.001 * ys + rng.normal(0, 1e-4, (H, W))


def test_median_fill_rounds():
    """The parallel-rounds hole fill: equal to the sequential fill
    (median_fill, the reference's order) on isolated holes; on blobs and
    border strips of a noisy plane it fills everything with values near the
    plane, as the sequential fill does; hole-free input is left as it is and
    an all-invalid frame stays invalid."""
    O = oracle()
    rng = np.random.default_rng(9)
    W, H = 64, 48
    d = _plane(W, H, rng)
    iso = d.copy()
    ys, xs = np.mgrid[1:H - 1:3, 1:W - 1:3]
    iso[ys.ravel(), xs.ravel()] = 0.0            # holes 3 apart: no two touch
    a = O.median_fill(iso)
    b, left = O.median_fill_rounds(iso)
    assert left == 0 and np.array_equal(a, b)
    blob = d.copy()
    blob[10:20, 10:25] = 0.0
    blob[:, :3] = 0.0
    blob[30:33, 40:60] = -1.0
    blob[rng.random((H, W)) < 0.05] = 0.0
    a = O.median_fill(blob)
    b, left = O.median_fill_rounds(blob)
    hole = blob <= 0
    assert left == 0 and (b > 0).all() and np.array_equal(b[~hole], blob[~hole])
    assert np.abs(b - d)[hole].max() < 0.05 and np.abs(a - d)[hole].max() < 0.05
    assert np.array_equal(b[hole], b[hole].astype(np.float32).astype(np.float64))   # medians are floats
    same, left = O.median_fill_rounds(d)
    assert left == 0 and np.array_equal(same, d)
    z, left = O.median_fill_rounds(np.zeros((H, W)))
    assert left == W * H and not z.any()
    # a 3 x 3 hole in a constant-gradient field fills ring by ring: the
    # centre (no known neighbour at first) comes last, from its 8 neighbours
    g = np.add.outer(np.arange(7.0), np.arange(7.0)) + 1.0
    h = g.copy()
    h[2:5, 2:5] = 0.0
    b, left = O.median_fill_rounds(h)
    assert left == 0 and b[3, 3] == np.float32(np.sort(b[2:5, 2:5].ravel()[[0, 1, 2, 3, 5, 6, 7, 8]])[4])


def test_colour_fill_rounds():
    """The parallel-rounds variant of the colour-guided small-hole fill:
    its exp against numpy's; equal colours give the plain mean of the known
    neighbours; across a colour edge the weights pick the hole's own side;
    regions of max_region pixels are filled, larger ones left as they are;
    a 3 x 3 hole fills its ring before its centre."""
    O = oracle()
    x = -np.concatenate([np.linspace(0, 10, 2001), [1e-12, 0.5 / 0.6, 3 / 0.6]])
    assert np.abs(O.fill_exp(x) / np.exp(x) - 1).max() < 4e-16
    W, H = 40, 30
    rng = np.random.default_rng(4)
    d = rng.uniform(1, 2, (H, W))
    col = np.full((H, W, 3), 0.5)
    h = d.copy()
    h[10, 10] = 0.0
    f, left = O.colour_fill_rounds(h, col)
    nb = np.delete(d[9:12, 9:12].ravel(), 4)
    assert left == 0 and abs(f[10, 10] - nb.mean()) < 1e-6 and np.array_equal(f[h > 0], d[h > 0])
    # colour edge at column 20: depth 1 on the left (dark), 2 on the right (bright)
    d2 = np.where(np.arange(W)[None, :] < 20, 1.0, 2.0) * np.ones((H, 1))
    col2 = np.where(np.arange(W)[None, :, None] < 20, 0.1, 0.9) * np.ones((H, 1, 3))
    h2 = d2.copy()
    h2[15, 19] = h2[5, 20] = 0.0
    f2, _ = O.colour_fill_rounds(h2, col2)
    w = np.exp(-3 * 0.8 ** 2 / 0.6)         # the other side's weight; one's own side weighs 1
    assert abs(f2[15, 19] - (5 * 1.0 + 3 * w * 2.0) / (5 + 3 * w)) < 1e-6
    assert abs(f2[5, 20] - (5 * 2.0 + 3 * w * 1.0) / (5 + 3 * w)) < 1e-6
    # a straight 4-connected run of 60 pixels is filled, one of 61 is not
    h3 = d.copy()
    h3[2, 0:30] = 0.0
    h3[3, 0:30] = 0.0            # 60 pixels
    h3[20, 0:31] = 0.0
    h3[21, 0:30] = 0.0           # 61 pixels
    f3, left = O.colour_fill_rounds(h3, col)
    assert left == 0 and (f3[2:4, 0:30] > 0).all() and (f3[20, 0:31] <= 0).all() and (f3[21, 0:30] <= 0).all()
    # the ring before the centre: the centre's neighbours are all filled values when it is filled
    h4 = d.copy()
    h4[5:8, 5:8] = 0.0
    f4, left = O.colour_fill_rounds(h4, col)
    assert left == 0 and abs(f4[6, 6] - np.delete(f4[5:8, 5:8].ravel(), 4).mean()) < 1e-6


def test_record_storage_rounding():
    """The 16-bit system records: BF16 equals torch's bfloat16 conversion of
    each f32 coefficient; FP16 scales the pixel's 13 coefficients by 2^-e so
    that the largest is in [2^14, 2^15), rounds with numpy's float16
    conversion and scales back -- values far below the largest lose their
    low bits (or flush to zero) as FP16's subnormals do.  A planted mistake
    (no scale: FP16 overflow above 65504) fails."""
    import torch
    rng = np.random.default_rng(8)
    rows = [rng.normal(size=13) * 10.0 ** rng.uniform(-8, 8) for _ in range(300)]
    rows += [np.r_[1e6, rng.normal(size=12) * 1e-6], np.r_[-3e-30, rng.normal(size=12) * 1e-31], np.zeros(13),
             np.r_[65504.0 * 3, np.ones(12)], np.r_[2.0 ** 14, 2.0 ** 15 - 1, np.full(11, 2.0 ** -11 * 3)]]
    c = np.array(rows, np.float32)
    got = O.record_round(c, O.REC_F16)
    m = np.abs(c).max(1)
    _, ex = np.frexp(m)
    e = np.where(m > 0, np.clip(ex - 1 - 14, -100, 100), 0)
    down = np.ldexp(np.float32(1), -e).astype(np.float32)[:, None]
    up = np.ldexp(np.float32(1), e).astype(np.float32)[:, None]
    want = ((c * down).astype(np.float16).astype(np.float32) * up).astype(np.float32)
    assert np.array_equal(got, want)
    big = (np.abs(got).max(1) > 0) & (np.abs(e) < 100)      # scale not clamped
    scaled = np.abs(got[big] * down[big]).max(1)
    assert np.all(scaled >= 2.0 ** 14) and np.all(scaled <= 2.0 ** 15)   # 2^15 - 1 rounds up to 2^15
    assert (np.abs(e) == 100).any()
    assert not np.array_equal(got, c.astype(np.float16).astype(np.float32))          # unscaled: overflows
    gotb = O.record_round(c, O.REC_BF16)
    assert np.array_equal(gotb, torch.from_numpy(c).to(torch.bfloat16).to(torch.float32).numpy())
    assert np.array_equal(O.record_round(c, O.REC_F32), c)
