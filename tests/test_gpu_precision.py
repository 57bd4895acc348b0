"""Precision variants: FP16 / BF16 storage of the per-pixel systems and the
relaxation factor omega.  omega = 1.8 is exact against the oracle run with the
same factor; the storage variants are the same computation on rounded
coefficients, so both schedules agree bit for bit, and the flow stays close to
the FP32 result (the distances are reported; the bound asserted is loose:
these variants change the reference's arithmetic on purpose)."""
import numpy as np
import pytest

from helpers import CONFIGS, hpb_for, load_package, make_case, oracle

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
    cfg = CONFIGS["cfg0_320x240"]
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
