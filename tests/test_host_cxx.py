"""The drop-in C++ API's host functions, driven through tests/cxx/test_host
(built with the library), on the CPU:

- recursiveMedianFilter (recursiveMedianFiltering.cpp:34-120), the fill the
  drop-in applies to both depth maps (sceneFlowMultiscale.cpp:54,61): bit for
  bit the oracle's sequential fill -- the same container and insert/erase
  sequence, hence the same visiting order -- on frames with isolated holes,
  blobs, border strips and mostly-invalid frames;
- recursiveMedianFilterSmallInvalidRegionsUsingColor (:127-222): its
  visiting order comes from boost's pair hash and is parity unpinned; what
  is order-independent is checked -- an isolated hole gets the colour-weighted
  mean of its known neighbours (the oracle's rounds variant agrees there),
  regions above the size limit stay invalid, hole-free frames are unchanged;
- readDepthMapValuesImg (main.cpp:32-48): millimetres times .001, rounded
  to float;
- downsizeFrame (rgbdFrameUtils.cpp:11-42, SMOOTH_DEPTH off): depth sampled
  at even rows and columns; the colour's 2x2 rounded mean stands for
  cv::resize's bilinear halving (OpenCV is absent: parity unpinned).
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, oracle

DRIVER = os.path.join(ROOT, "tests", "cxx", "build", "test_host")


def _run(op, payload, tmp_path):
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is not built; run __graft_entry__.build()")
    src, dst = tmp_path / f"{op}.in", tmp_path / f"{op}.out"
    src.write_bytes(payload)
    r = subprocess.run([DRIVER, op, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    return r.returncode, (dst.read_bytes() if dst.exists() else b""), r.stderr


def _frame(W, H, seed, pattern):
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    d = 1.5 + 0.002 * xs - 0.001 * ys + rng.normal(0, 1e-3, (H, W))
    if pattern == "isolated":
        yy, xx = np.mgrid[1:H - 1:3, 1:W - 1:3]
        d[yy.ravel(), xx.ravel()] = 0.0
    elif pattern == "blobs":
        d[rng.random((H, W)) < 0.05] = 0.0
        d[H // 4:H // 4 + H // 8, W // 5:W // 5 + W // 6] = 0.0
        d[:, :3] = -1.0
        d[H - 2:, :] = 0.0
    elif pattern == "sparse":
        d[rng.random((H, W)) < 0.85] = 0.0
    return d


@pytest.mark.parametrize("W,H,pattern", [(64, 48, "isolated"), (160, 120, "blobs"), (97, 61, "sparse"),
                                         (64, 48, "none")])
def test_median_fill_equals_the_oracle(tmp_path, W, H, pattern):
    d = _frame(W, H, W * 7 + H, pattern)
    rc, out, err = _run("fill", np.array([W, H], np.int32).tobytes() + d.tobytes(), tmp_path)
    assert rc == 0, err
    got = np.frombuffer(out, np.float64).reshape(H, W)
    want = oracle().median_fill(d)
    assert np.array_equal(got, want)
    assert (got > 0).all()


def test_median_fill_refuses_an_all_invalid_frame(tmp_path):
    d = np.zeros((12, 16))
    rc, _, err = _run("fill", np.array([16, 12], np.int32).tobytes() + d.tobytes(), tmp_path)
    assert rc == 1 and "no valid neighbour" in err


def _colour_fill(tmp_path, d, bgr, max_region=60):
    H, W = d.shape
    rc, out, err = _run("colourfill", np.array([W, H, max_region], np.int32).tobytes() + d.tobytes() +
                        np.ascontiguousarray(bgr, np.float64).tobytes(), tmp_path)
    assert rc == 0, err
    return np.frombuffer(out, np.float64).reshape(H, W)


def test_colour_fill_isolated_holes_are_colour_weighted_means(tmp_path):
    O = oracle()
    W, H = 40, 30
    rng = np.random.default_rng(4)
    d = rng.uniform(1, 2, (H, W))
    col = rng.uniform(0, 1, (H, W, 3))
    h = d.copy()
    yy, xx = np.mgrid[1:H - 1:3, 1:W - 1:3]
    h[yy.ravel(), xx.ravel()] = 0.0
    got = _colour_fill(tmp_path, h, col)
    for (i, j) in zip(yy.ravel(), xx.ravel()):
        nb = [(a, b) for a in (i - 1, i, i + 1) for b in (j - 1, j, j + 1) if (a, b) != (i, j)]
        wts = np.array([np.exp(-np.sum((col[a, b] - col[i, j]) ** 2) / 0.6) for a, b in nb])
        want = np.sum(wts * np.array([d[a, b] for a, b in nb])) / wts.sum()
        assert abs(got[i, j] - want) < 1e-6 * want
    assert np.array_equal(got[h > 0], d[h > 0])
    rounds, left = O.colour_fill_rounds(h, col)
    assert left == 0 and np.allclose(got, rounds, rtol=1e-6, atol=0)


def test_colour_fill_leaves_large_regions_and_hole_free_frames(tmp_path):
    W, H = 40, 30
    rng = np.random.default_rng(5)
    d = rng.uniform(1, 2, (H, W))
    col = rng.uniform(0, 1, (H, W, 3))
    assert np.array_equal(_colour_fill(tmp_path, d, col), d)
    h = d.copy()
    h[5:14, 5:14] = 0.0          # 81 pixels: above the limit
    h[20:22, 10:15] = 0.0        # 10 pixels: filled
    got = _colour_fill(tmp_path, h, col)
    assert (got[5:14, 5:14] <= 0).all() and (got[20:22, 10:15] > 0).all()
    assert np.array_equal(got[h > 0], d[h > 0])


def test_read_depth_map_values(tmp_path):
    W, H = 7, 5
    mm = np.arange(W * H, dtype=np.uint16).reshape(H, W) * 1873
    mm[0, 0], mm[0, 1], mm[0, 2] = 0, 1, 65535
    rc, out, err = _run("readdepth", np.array([W, H], np.int32).tobytes() + mm.tobytes(), tmp_path)
    assert rc == 0, err
    got = np.frombuffer(out, np.float32).reshape(H, W)
    assert np.array_equal(got, (mm.astype(np.int64) * .001).astype(np.float32))
    assert got[0, 0] == 0 and got[0, 2] == np.float32(65.535)


@pytest.mark.parametrize("W,H", [(8, 6), (9, 7)])
def test_downsize_frame(tmp_path, W, H):
    rng = np.random.default_rng(W)
    img = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    depth = rng.uniform(0.5, 4, (H, W)).astype(np.float32)
    rc, out, err = _run("downsize", np.array([W, H], np.int32).tobytes() + img.tobytes() + depth.tobytes(), tmp_path)
    assert rc == 0, err
    w, h = np.frombuffer(out[:8], np.int32)
    assert (w, h) == (W // 2, H // 2)
    small = np.frombuffer(out[8:8 + 3 * w * h], np.uint8).reshape(h, w, 3)
    dsmall = np.frombuffer(out[8 + 3 * w * h:], np.float32).reshape(h, w)
    assert np.array_equal(dsmall, depth[0:2 * h:2, 0:2 * w:2])
    s = img[0:2 * h:2, 0:2 * w:2].astype(int) + img[0:2 * h:2, 1:2 * w:2] + img[1:2 * h:2, 0:2 * w:2] + \
        img[1:2 * h:2, 1:2 * w:2]
    assert np.array_equal(small, ((s + 2) >> 2).astype(np.uint8))
