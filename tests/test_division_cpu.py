"""The tiled relaxation divides by a per-pixel reciprocal with one Markstein
correction (q0 = a*y, r = fma(-b, q0, a), q = fma(r, y, q0), y = 1/b) inside a
guarded domain, and must give the IEEE quotient bit for bit.  Checked on the
CPU against a / b over sampled operands, with a negative control (the
uncorrected product) that the check must catch.  A larger offline run of the
same checker (1.5e9 samples per mode, 6e9 in all) found no mismatch."""
import pytest

from helpers import oracle


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_reciprocal_division_is_ieee(mode):
    bad, first = oracle().check_reciprocal_division(3_000_000, 100 + mode, mode)
    assert bad == 0, first


@pytest.mark.parametrize("mode", [0, 2])
def test_negative_control_is_caught(mode):
    bad, _ = oracle().check_reciprocal_division(200_000, 7, mode, recipe=1)
    assert bad > 1000


@pytest.mark.parametrize("d", [12, 5, 3])
def test_division_by_constant_is_ieee(d):
    """The kernels' f64 divisions by 12 (derivative stencils), 5 (channel
    means) and 3 (Laplacian weight means): Markstein's correction is bit-equal
    to IEEE x / d on the sampled domain; the bare product is not."""
    bad, x = oracle().check_div_const(5_000_000, 3 + d, d)
    assert bad == 0, x
    bad, _ = oracle().check_div_const(200_000, 4 + d, d, recipe=1)
    assert bad > 1000


@pytest.mark.parametrize("d", [12, 5, 3])
def test_f32_division_by_constant_is_ieee(d):
    """The f32 stage of the 16-bit image variants: the same recipe in f32 is
    bit-equal to IEEE x / d for every significand of a binade (which the
    power-of-two scaling carries to the whole domain [2^-100, 2^100]); the
    bare product is not."""
    bad, x = oracle().check_div_const_f32(d)
    assert bad == 0, x
    bad, _ = oracle().check_div_const_f32(d, recipe=1)
    assert bad > 1000


@pytest.mark.parametrize("bits", [11, 8])
def test_refined_reciprocal_of_short_significands_is_ieee(bits):
    """The kernels' reciprocals of a 16-bit record's decoded diagonal
    (v_rcp_f32 within one ulp, then one Newton step with FMAs): for every
    significand of 11 (FP16) or 8 (BF16) significant bits and every starting
    value within one ulp, the IEEE quotient 1.0f / c bit for bit.  The
    unrefined start (negative control) is caught, and so is the same
    refinement of full 24-bit significands, where the method's premise (1/c
    far from a rounding midpoint) does not hold."""
    bad, c = oracle().check_recip_short(bits)
    assert bad == 0, c
    bad, _ = oracle().check_recip_short(bits, recipe=1)
    assert bad > 100
    bad, _ = oracle().check_recip_short(bits, recipe=2)
    assert bad > 0
