"""Precision variants: FP16 / BF16 storage of the per-pixel systems and the
relaxation factor omega.  omega = 1.8 is exact against the oracle run with the
same factor; the storage variants are the same computation on rounded
coefficients, so both schedules agree bit for bit, and the flow stays close to
the FP32 result (the distances are reported; the bound asserted is loose:
these variants change the reference's arithmetic on purpose)."""
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


def _case(W=160, H=120, seed=61, mw=20):
    pair, im1, d1, im2, d2, sigma = make_case(W, H, seed, 4)
    hpb, lv = hpb_for(pair, W, H, 0.5, mw)
    return pair, im1, d1, im2, d2, sigma, hpb


@pytest.mark.parametrize("variant", ["plain", "tiled"])
def test_omega_matches_oracle(ctx, R, variant):
    O = oracle()
    W, H = 160, 120
    pair, im1, d1, im2, d2, sigma, hpb = _case(W, H)
    kw = dict(n_outer=4, n_sor=30, min_width=20)
    ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb, O.make_params(W, H, omega=1.8, **kw))
    pg = R.make_params(W, H, omega=1.8, relax_variant=R.RELAX_TILED if variant == "tiled" else R.RELAX_PLAIN, **kw)
    ub, vb, wb, sb = ctx.multiscale(pg, im1, d1, im2, d2, sigma, hpb)
    assert [sb.sweeps_exec[k] for k in range(sa.n_levels)] == [sa.sweeps_exec[k] for k in range(sa.n_levels)]
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_storage_variants(ctx, R, storage):
    O = oracle()
    cfg = WORKLOADS["320x240_10x30"]
    W, H = cfg["W"], cfg["H"]
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 11, cfg["segments"], cfg["f"], cfg["cx"], cfg["cy"])
    hpb, lv = hpb_for(pair, W, H, 0.5, cfg["min_width"])
    kw = dict(f=cfg["f"], cx=cfg["cx"], cy=cfg["cy"], n_outer=cfg["n_outer"], n_sor=cfg["n_sor"],
              min_width=cfg["min_width"])
    st = R.STORE_F16 if storage == "f16" else R.STORE_BF16
    ref = ctx.multiscale(R.make_params(W, H, relax_variant=R.RELAX_TILED, **kw), im1, d1, im2, d2, sigma, hpb)
    tiled = ctx.multiscale(R.make_params(W, H, relax_variant=R.RELAX_TILED, storage=st, **kw), im1, d1, im2, d2,
                           sigma, hpb)
    plain = ctx.multiscale(R.make_params(W, H, relax_variant=R.RELAX_PLAIN, storage=st, **kw), im1, d1, im2, d2,
                           sigma, hpb)
    for a, b in zip(tiled[:3], plain[:3]):
        assert np.array_equal(a, b)          # both schedules, same rounded systems
    epe = np.sqrt((tiled[0] - ref[0]) ** 2 + (tiled[1] - ref[1]) ** 2)
    uf, vf, wf, _ = O.multiscale(im1, d1, im2, d2, sigma, hpb, O.make_params(W, H, precision=O.PREC_F64, **kw))
    epe64 = np.sqrt((tiled[0] - uf) ** 2 + (tiled[1] - vf) ** 2)
    epe32 = np.sqrt((ref[0] - uf) ** 2 + (ref[1] - vf) ** 2)
    print(f"{storage}: EPE vs FP32 mean {epe.mean():.4g} median {np.median(epe):.4g} p99 {np.percentile(epe, 99):.4g}; "
          f"vs FP64 median {np.median(epe64):.4g} (FP32 vs FP64 median {np.median(epe32):.4g}); "
          f"|dw| vs FP32 median {np.median(np.abs(tiled[2] - ref[2])):.3g} m")
    assert np.median(epe) < 0.05 and np.isfinite(tiled[0]).all()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_record_storage_matches_oracle(ctx, R, storage):
    """16-bit system records: the device equals the oracle run with the same
    record rounding (FP16 with a per-pixel power-of-two scale, or BF16),
    bit for bit and with the same iterations, in both relaxation schedules;
    with the point-SOR order too.  The block update refuses them."""
    O = oracle()
    W, H = 160, 120
    pair, im1, d1, im2, d2, sigma, hpb = _case(W, H, seed=63)
    st = R.STORE_F16 if storage == "f16" else R.STORE_BF16
    with pytest.raises(R.RgbdFlowError) as e:     # the block update needs f32 records
        ctx.multiscale(R.make_params(W, H, storage=st, pixel_update=R.PIXEL_BLOCK, omega=1.8), im1, d1, im2, d2,
                       sigma, hpb)
    assert e.value.code == R.ERR_UNSUPPORTED
    for mode in (R.PIXEL_JACOBI, R.PIXEL_GS):
        kw = dict(n_outer=4, n_sor=40, min_width=20, omega=0.7 if mode == R.PIXEL_JACOBI else 1.8)
        ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb,
                                      O.make_params(W, H, record_storage=st, pixel_gs=mode, **kw))
        for variant in (R.RELAX_PLAIN, R.RELAX_TILED):
            print(f"{storage} records, pixel update {mode}, relax variant {variant}", flush=True)
            pg = R.make_params(W, H, relax_variant=variant, storage=st, pixel_update=mode, **kw)
            ub, vb, wb, sb = ctx.multiscale(pg, im1, d1, im2, d2, sigma, hpb)
            L = sa.n_levels
            assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
            assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


# ----------------------------------------- point SOR order and the FP64 residual
@pytest.fixture(scope="module")
def system():
    O = oracle()
    from helpers import level_inputs
    li, (u, v, w), _ = level_inputs(seed=8, W=96, H=72)
    return O.level_system(li, O.make_params(96, 72), u, v, w)


SCHEDULES = {
    "plain": dict(relax_variant=0),
    "tiled": dict(relax_variant=1, tile_iters=5, tile_w=24, tile_h=16),
    "tiled_single": dict(relax_variant=1),
}


@pytest.mark.parametrize("sched", list(SCHEDULES))
@pytest.mark.parametrize("mode,omega", [(0, 0.7), (1, 1.8), (1, 1.0), (0, 1.3), (2, 1.8), (2, 1.0)])
def test_relax_residual_matches_oracle(ctx, R, system, sched, mode, omega):
    """The FP64 residual exit on the device: the same iterations, the same
    increments bit for bit, and residual / energy equal to the oracle's f64
    reduction up to summation order.  mode: the pixel update (0 the
    reference's, 1 point SOR, 2 the 3x3 block solve)."""
    O = oracle()
    x0 = np.zeros((3, 72, 96))
    xa, ia, ra, ea = O.relax_residual(system, 600, 1e-4, x0, omega=omega, pixel_gs=mode)
    p = R.make_params(96, 72, omega=omega, pixel_update=mode, **SCHEDULES[sched])
    xb, ib, rb, eb = ctx.relax_residual(system, 600, 1e-4, x0, p)
    print(f"{sched} mode={mode} omega={omega}: {ib} iterations, residual {rb:.3e}, energy {eb:.9g}")
    assert ia == ib
    assert np.array_equal(xa, xb), float(np.abs(xa - xb).max())
    assert abs(ra - rb) <= 1e-10 * ra and abs(ea - eb) <= 1e-10 * abs(ea)


@pytest.mark.parametrize("sched", list(SCHEDULES))
def test_block_update_non_definite_pixels_match_oracle(ctx, R, system, sched):
    """Pixels whose block is not negative definite (Buu made positive at every
    37th pixel) keep their values in the block update, on the device as in
    the oracle: bit for bit, with the same iterations."""
    O = oracle()
    c = system.copy()
    c[4].ravel()[::37] = np.abs(c[4].ravel()[::37])
    x0 = np.zeros((3, 72, 96))
    xa, ia, ra, ea = O.relax_residual(c, 100, 0.0, x0, omega=1.8, pixel_gs=2)
    xb, ib, rb, eb = ctx.relax_residual(c, 100, 0.0, x0, R.make_params(96, 72, omega=1.8, pixel_update=2,
                                                                       **SCHEDULES[sched]))
    assert ia == ib == 100
    assert np.array_equal(xa, xb), float(np.abs(xa - xb).max())
    assert np.isfinite(xb).all()


@pytest.mark.parametrize("mode", [1, 2])
def test_sor_converged_energy_agrees(ctx, R, system, mode):
    """Red-black SOR with omega = 1.8 -- point (mode 1) or 3x3 block (mode 2)
    -- reaches the reference iteration's converged energy (omega = 0.7, old
    values within the pixel) within 1e-4 relative, in far fewer iterations
    (device residuals and energies).  The f32 increments floor the relative
    residual near 5e-6 here, so the target is 2e-5."""
    x0 = np.zeros((3, 72, 96))
    tiled = dict(relax_variant=R.RELAX_TILED)
    _, i_ref, r_ref, e_ref = ctx.relax_residual(system, 40000, 2e-5, x0, R.make_params(96, 72, **tiled))
    _, i_sor, r_sor, e_sor = ctx.relax_residual(system, 40000, 2e-5, x0,
                                                R.make_params(96, 72, omega=1.8, pixel_update=mode, **tiled))
    print(f"reference order omega 0.7: {i_ref} its, energy {e_ref:.12g}; mode {mode} 1.8: {i_sor} its, {e_sor:.12g}")
    assert r_ref <= 2e-5 and r_sor <= 2e-5
    assert abs(e_sor - e_ref) <= 1e-4 * abs(e_ref)
    assert i_sor * 3 < i_ref


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("variant", ["plain", "tiled"])
def test_point_sor_pair_matches_oracle(ctx, R, variant, mode):
    """Whole pairs with the point SOR order or the 3x3 block solve (omega
    1.8): bit-exact with the oracle, and with the FP64 residual exit the same
    iterations per level."""
    O = oracle()
    W, H = 160, 120
    pair, im1, d1, im2, d2, sigma, hpb = _case(W, H)
    kw = dict(n_outer=4, n_sor=60, min_width=20, omega=1.8)
    rv = R.RELAX_TILED if variant == "tiled" else R.RELAX_PLAIN
    for resid in (False, True):
        po = O.make_params(W, H, pixel_gs=mode, residual_exit=resid, residual_tol=1e-3, **kw)
        pg = R.make_params(W, H, pixel_update=mode, relax_variant=rv, residual_tol=1e-3,
                           inner_exit=R.EXIT_RESIDUAL if resid else R.EXIT_REFERENCE, **kw)
        ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb, po)
        ub, vb, wb, sb = ctx.multiscale(pg, im1, d1, im2, d2, sigma, hpb)
        L = sa.n_levels
        assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
        assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)
        if resid:
            assert [sb.residual_checks[k] for k in range(L)] == [sa.residual_checks[k] for k in range(L)]
            for k in range(L):
                assert abs(sb.residual[k] - sa.residual[k]) <= 1e-10 * sa.residual[k]
                assert abs(sb.energy[k] - sa.energy[k]) <= 1e-10 * abs(sa.energy[k])


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_relax_storage_schedules_agree(ctx, R, system, storage):
    """The relaxation stage entry re-stores the records as the level solve
    does: both schedules give the same bits on the rounded system, and the
    result stays near the f32 one."""
    st = R.STORE_F16 if storage == "f16" else R.STORE_BF16
    x0 = np.zeros((3, 72, 96))
    a = ctx.relax(system, 30, x0, R.make_params(96, 72, relax_variant=R.RELAX_PLAIN, storage=st))
    b = ctx.relax(system, 30, x0, R.make_params(96, 72, relax_variant=R.RELAX_TILED, tile_iters=5, tile_w=24,
                                                tile_h=16, storage=st))
    f = ctx.relax(system, 30, x0, R.make_params(96, 72, relax_variant=R.RELAX_PLAIN))
    assert np.array_equal(a, b)
    assert not np.array_equal(a, f)
    # bf16 keeps 8 significand bits (relative rounding 2^-9), f16 with its scale 11
    rel = 1e-3 if storage == "f16" else 1e-2
    assert np.median(np.abs(a - f)) < rel * np.median(np.abs(f)) + 1e-7


# ------------------------------------------------ FP16 / BF16 feature images
@pytest.mark.parametrize("img", ["f16", "bf16"])
def test_image_storage_matches_oracle(ctx, R, img):
    """Feature images (frame 1, frame 2, warped frame 2) stored in FP16 /
    BF16: the device equals the oracle run with the same rounding, bit for
    bit, with both relaxation schedules and as sceneFlowWithInit; the flow
    stays close to the f64-image result (distances reported, loose bound:
    the variant changes the reference's arithmetic on purpose).  The
    single-level entry points, which take f64 images, refuse the variant."""
    O = oracle()
    W, H = 160, 120
    pair, im1, d1, im2, d2, sigma, hpb = _case(W, H)
    kw = dict(n_outer=4, n_sor=30, min_width=20)
    st = R.IMAGES_F16 if img == "f16" else R.IMAGES_BF16
    ua, va, wa, sa = O.multiscale(im1, d1, im2, d2, sigma, hpb, O.make_params(W, H, image_storage=st, **kw))
    for variant in (R.RELAX_PLAIN, R.RELAX_TILED):
        print(f"{img} images, relax variant {variant}", flush=True)
        ub, vb, wb, sb = ctx.multiscale(R.make_params(W, H, relax_variant=variant, image_storage=st, **kw),
                                        im1, d1, im2, d2, sigma, hpb)
        assert [sb.sweeps_exec[k] for k in range(sa.n_levels)] == [sa.sweeps_exec[k] for k in range(sa.n_levels)]
        assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)
    ref = ctx.multiscale(R.make_params(W, H, relax_variant=R.RELAX_TILED, **kw), im1, d1, im2, d2, sigma, hpb)
    epe = np.sqrt((ub - ref[0]) ** 2 + (vb - ref[1]) ** 2)
    uf, vf, wf, _ = O.multiscale(im1, d1, im2, d2, sigma, hpb, O.make_params(W, H, precision=O.PREC_F64, **kw))
    epe64 = np.sqrt((ub - uf) ** 2 + (vb - vf) ** 2)
    print(f"{img} images: EPE vs f64 images median {np.median(epe):.4g} p99 {np.percentile(epe, 99):.4g}; "
          f"vs FP64 oracle median {np.median(epe64):.4g}; |dw| median {np.median(np.abs(wb - ref[2])):.3g} m")
    assert np.median(epe) < 0.05 and np.isfinite(ub).all()
    # sceneFlowWithInit from an estimate
    est = pair["flow"].astype(np.float64) * 0.7
    po = O.make_params(W, H, n_outer=3, n_sor=30, image_storage=st)
    ua, va, wa, _ = O.with_init(im1, d1, im2, d2, sigma, None, est, po)
    pw = R.make_params(W, H, n_outer=3, n_sor=30, relax_variant=R.RELAX_TILED, image_storage=st)
    ub, vb, wb, _ = ctx.with_init(pw, im1, d1, im2, d2, est, sigma, None)
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)
    li, (u, v, w), _ = level_inputs(seed=9, W=96, H=72)
    with pytest.raises(R.RgbdFlowError) as e:
        ctx.solve_level(li, R.make_params(96, 72, image_storage=st), True, u, v, w)
    assert e.value.code == R.ERR_UNSUPPORTED
