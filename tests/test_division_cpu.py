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


def test_div12_is_ieee():
    """The derivative stencils' f64 division by 12 (kernels_solve.hip div12)."""
    bad, x = oracle().check_div12(5_000_000, 3)
    assert bad == 0, x
    bad, _ = oracle().check_div12(200_000, 4, recipe=1)
    assert bad > 1000
