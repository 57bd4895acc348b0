"""Pins of the oracle's f32 stage (the 16-bit image variants: every per-outer
stage formula in f32, not in the reference).  On the same level inputs --
feature images already representable in FP16, so the variants' rounding of
them is the identity -- the f32 stage must reproduce the reference-precision
stage (f64 with its f32 roundings; itself pinned in test_oracle_pins*.py) to
f32 accuracy: a dropped term, a swapped weight or a wrong sign would move a
coefficient by far more than f32 rounding does.  And it must not be the f64
stage itself (the variant path is taken)."""
import numpy as np

from helpers import level_inputs, oracle

O = oracle()


def _f16_inputs(seed, W, H):
    li, (u, v, w), pair = level_inputs(seed=seed, W=W, H=H)
    r16 = lambda a: O.img_round(a, O.IMG_F16)
    li16 = O.LevelInputs(r16(li.I1), r16(li.I2), r16(li.warpI2), li.D1, li.D2, li.warpD2, li.delta, li.s, li.m,
                         li.r, li.focal_level)
    return li16, (u, v, w)


def test_f32_stage_system_is_the_reference_system_to_f32_accuracy():
    W, H = 48, 36
    li, (u, v, w) = _f16_inputs(31, W, H)
    ref = O.level_system(li, O.make_params(W, H), u, v, w)
    f32 = O.level_system(li, O.make_params(W, H, image_storage=O.IMG_F16), u, v, w)
    assert not np.array_equal(ref, f32)
    # the depth term's weight 1/2 / |Dt + ...| comes from a difference of depths of ~1-4 m, whose
    # f32 cancellation leaves ~3e-7 m: up to ~3e-3 of the largest coefficient in the planes it
    # enters (Bww, Buw, Bvw, rw); the others stay within ~1e-4
    depth_planes = (6, 8, 9, 12)
    for k in range(13):
        scale = np.abs(ref[k]).max()
        tol = 1e-2 if k in depth_planes else 5e-4
        assert np.abs(f32[k] - ref[k]).max() <= tol * scale + 1e-30, k
    # and the BF16 image variant runs the same f32 stage
    b = O.level_system(li, O.make_params(W, H, image_storage=O.IMG_BF16), u, v, w)
    assert np.array_equal(b, f32)


def test_f32_stage_weights_are_the_reference_weights_to_f32_accuracy():
    W, H = 48, 36
    li, (u, v, w) = _f16_inputs(32, W, H)
    ref = O.level_weights(li, O.make_params(W, H), u, v, w)
    f32 = O.level_weights(li, O.make_params(W, H, image_storage=O.IMG_F16), u, v, w)
    phi, psi, pi = ref
    phi32, psi32, pi32 = f32
    assert np.allclose(phi32, phi, rtol=1e-4, atol=0)
    assert np.allclose(psi32, psi, rtol=1e-5, atol=0)
    assert np.allclose(pi32, pi, rtol=2e-2, atol=0)      # the depth difference's f32 cancellation
    assert not all(np.array_equal(a, b) for a, b in zip(ref, f32))


def test_f32_stage_warp_is_the_reference_warp_rounded():
    W, H = 40, 30
    li, (u, v, w) = _f16_inputs(33, W, H)
    wI, wD = O.warp_gpu_style(li.I1, li.D1, li.I2, li.D2, u, v, w)
    fI, fD = O.warp_gpu_style_f32(li.I1, li.D1, li.I2, li.D2, u, v, w, O.IMG_F16)
    # the warped features end in FP16: within one FP16 step of the reference's, rounded
    rI = O.img_round(wI, O.IMG_F16)
    assert np.all(np.abs(fI - rI) <= np.abs(rI) * 2.0 ** -10 + 2.0 ** -24)
    assert np.mean(fI == rI) > 0.95
    # the warped depth stays f64 storage: f32 sums of f32-rounded depths
    assert np.allclose(fD, wD, rtol=1e-6, atol=1e-6)
    assert not np.array_equal(fD, wD)
