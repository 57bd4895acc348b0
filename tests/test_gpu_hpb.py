"""The boundary probabilities on the device (kernels_hpb.hip, the C ABI's
rgbdflow_heuristic_pb) against the oracle's restatement (oracle/hpb.cpp):
bit for bit, on the closed-form scenes the oracle pins use and on synthetic
frames at the benchmark's sizes (levels of a 640x480 pyramid included)."""
import numpy as np
import pytest

from helpers import load_package, make_case, oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    R = load_package()
    c = R.Context(0)
    yield c
    c.close()


def _scenes():
    W, H = 64, 48
    out = []
    step = np.where(np.arange(W)[None, :] < 30, 1.0, 1.5) * np.ones((H, 1))
    out.append(("depth_step", np.full((H, W, 3), 0.5), step))
    colour = np.where(np.arange(W)[None, :, None] < 30, 0.0, 1.0) * np.ones((H, 1, 3))
    out.append(("colour_step_no_depth", colour, np.zeros((H, W))))
    hole = np.full((H, W), 1.2)
    hole[20:30, 25:35] = 0
    hole[:, :6] = 0
    out.append(("holes", np.full((H, W, 3), 0.5), hole))
    return out


@pytest.mark.parametrize("k", range(3))
def test_closed_form_scenes(ctx, k):
    O = oracle()
    name, bgr, depth = _scenes()[k]
    print(f"hpb scene {name}", flush=True)
    a = O.heuristic_pb(bgr, depth)
    b = ctx.heuristic_pb(bgr, depth)
    assert np.array_equal(a, b), (name, float(np.abs(a - b).max()))


@pytest.mark.parametrize("W,H,seed", [(160, 120, 3), (320, 240, 4), (640, 480, 5), (101, 77, 6)])
def test_synthetic_frames(ctx, W, H, seed):
    O = oracle()
    pair, im1, d1, im2, d2, sigma = make_case(W, H, seed, 6)
    d = d1.copy()
    d[H // 3:H // 3 + 7, W // 4:W // 4 + 9] = 0            # a hole
    print(f"hpb frame {W}x{H} seed {seed}", flush=True)
    a = O.heuristic_pb(im1, d)
    b = ctx.heuristic_pb(im1, d)
    diff = a != b
    assert not diff.any(), (int(diff.sum()), float(np.abs(a - b).max()))
    assert a.max() > 0.5 and (a > 0).mean() < 0.3            # boundaries found, most edges not boundaries
