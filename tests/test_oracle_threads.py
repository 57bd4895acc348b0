"""The oracle's level solve on row partitions (src/parallelism.cpp:24-33
partitionEvenly) gives the same bits as the scalar restatement for any thread
count: whole pairs, both precisions, with and without the early exits, and
the bare relaxation.  The all-cores CPU baseline relies on this."""
import numpy as np
import pytest

from helpers import hpb_for, make_case, oracle


@pytest.fixture
def O():
    O = oracle()
    yield O
    O.set_threads(1)


def _pair(O, precision, exits, W=120, H=90):
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 17, 4)
    hpb, _ = hpb_for(pair, W, H, 0.5, 30)
    p = O.make_params(W, H, n_outer=4, n_sor=30, min_width=30, precision=precision,
                      disable_inner_exit=not exits, disable_outer_exit=not exits)
    u, v, w, st = O.multiscale(im1, d1, im2, d2, sigma, hpb, p)
    return u, v, w, [st.sweeps_exec[k] for k in range(st.n_levels)]


@pytest.mark.parametrize("precision,exits", [(0, True), (0, False), (1, True)])
def test_threads_do_not_change_whole_pairs(O, precision, exits):
    prec = [O.PREC_REF, O.PREC_F64][precision]
    O.set_threads(1)
    ref = _pair(O, prec, exits)
    for n in (3, 8):
        O.set_threads(n)
        assert O.threads() == n
        got = _pair(O, prec, exits)
        for a, b in zip(ref[:3], got[:3]):
            assert np.array_equal(a, b)
        assert ref[3] == got[3]


def test_threads_do_not_change_relaxation(O):
    rng = np.random.default_rng(5)
    W, H = 61, 47   # odd sizes: uneven partitions and colour phases
    coef = np.zeros((13, H, W))
    coef[0:4] = rng.uniform(0, .2, (4, H, W))
    coef[4:7] = -rng.uniform(1, 2, (3, H, W))
    coef[7:10] = rng.uniform(-.05, .05, (3, H, W))
    coef[10:13] = rng.uniform(-.1, .1, (3, H, W))
    x0 = rng.uniform(-.05, .05, (3, H, W))
    O.set_threads(1)
    a = O.relax(coef, 20, x0.copy())
    O.set_threads(5)
    b = O.relax(coef, 20, x0.copy())
    assert np.array_equal(a, b)
