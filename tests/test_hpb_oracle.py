"""Pins of the oracle's boundary-probability restatement (oracle/hpb.cpp,
computeHeuristicPb of src/heuristicBoundaries.cpp:40-358) by closed forms:
the exp it uses, normals of exact planes, and the probability of depth
steps, colour steps over invalid depth, holes (the invalid-run walk with the
reference's direction codes), the image-edge hack and non-maximum
suppression."""
import numpy as np
import pytest

from helpers import oracle

O = oracle()


def plane_depth(W, H, n, d, f=None):
    """Depth of the plane n.X = d seen by the HPB camera (f = 525 W / 640,
    c = (W/2, H/2) with integer halves)."""
    f = np.float32(525.0 * W / 640) if f is None else f
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    rx, ry = (xs - (W // 2)) / f, (ys - (H // 2)) / f
    return d / (n[0] * rx + n[1] * ry + n[2])


def test_exp_is_correctly_rounded_almost_everywhere():
    rng = np.random.default_rng(1)
    x = np.concatenate([-rng.uniform(0, 30, 20000), -rng.uniform(0, 1e-3, 2000), [0.0, -1e-30, -87.0, -103.0]])
    x = x.astype(np.float32)
    got = O.hpb_exp(x).astype(np.float64)
    exact = np.exp(x.astype(np.float64))
    want = exact.astype(np.float32).astype(np.float64)
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - exact) <= ulp)            # within one float ulp everywhere
    assert np.mean(got == want) > 0.999                  # and correctly rounded almost always
    assert O.hpb_exp([0.0])[0] == 1.0 and O.hpb_exp([-200.0])[0] == 0.0


def test_smallest_eigenvector_matches_eigh():
    """The normals' eigen-solver against numpy's eigh on random symmetric PSD
    matrices shaped like point covariances (two spreads, one small normal
    spread, eigenvalue ratio l0 / l1 up to 0.2), plus the rank-2 case; a
    planted mistake (the dominant instead of the smallest eigenvector)
    fails."""
    rng = np.random.default_rng(3)
    worst = 0.0
    for t in range(400):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        l1, l2 = rng.uniform(1e-4, 1e-2, 2)
        l0 = rng.uniform(0, 0.2) * min(l1, l2) if t % 10 else 0.0
        a = (q * np.array([l0, l1, l2])) @ q.T
        a = (a + a.T) / 2
        n = O.hpb_smallest_eigvec(a)
        w, v = np.linalg.eigh(a)
        ang = np.arccos(min(1.0, abs(float(n @ v[:, 0])) / np.linalg.norm(n)))
        worst = max(worst, ang)
        assert abs(np.linalg.norm(n) - 1) < 1e-12
        assert np.arccos(min(1.0, abs(float(n @ v[:, 2])))) > 0.1   # not the dominant one
    assert worst < 1e-6, worst
    assert not O.hpb_smallest_eigvec(np.zeros((3, 3))).any()


@pytest.mark.parametrize("n", [(0, 0, -1.0), (0.3, -0.2, -0.93), (-0.5, 0.1, -0.86)])
def test_normals_of_a_plane(n):
    W, H = 64, 48
    n = np.array(n) / np.linalg.norm(n)
    depth = -plane_depth(W, H, n, 1.5)                   # camera-facing: n points at the camera
    assert (depth > 0.5).all() and (depth < 4).all()
    nrm, val = O.hpb_normals(depth)
    corners = np.zeros((H, W), bool)
    corners[[0, 0, -1, -1], [0, -1, 0, -1]] = True       # 5 x 5 windows: fewer than 27 points
    assert (val == ~corners).all()
    assert not nrm[corners].any()
    assert np.abs(nrm[~corners] - n.astype(np.float32)).max() < 2e-4


def test_flat_uniform_scene_has_no_boundaries():
    W, H = 64, 48
    bgr = np.full((H, W, 3), 0.5)
    depth = np.full((H, W), 1.7)
    assert not O.heuristic_pb(bgr, depth).any()


def test_depth_step_is_one_edge():
    W, H, c = 64, 48, 30
    bgr = np.full((H, W, 3), 0.5)
    depth = np.where(np.arange(W)[None, :] < c, 1.0, 1.5) * np.ones((H, 1))
    p = O.heuristic_pb(bgr, depth)
    assert np.all(p[:, c - 1, 0] == 1.0)                 # dz / ratio far above 0.02: 1 - exp(-huge)
    p2 = p.copy()
    p2[:, c - 1, 0] = 0
    assert np.abs(p2).max() < 1e-3                       # no other edge survives


def test_colour_step_over_invalid_depth():
    """Without depth the colour half-disks decide: radius 5, black | white in
    every channel gives a mean difference of 1 per channel, so
    p = 1 - exp(-(sqrt(3) / 0.7)^2) at the step, suppressed beside it."""
    W, H, c = 64, 48, 30
    bgr = np.where(np.arange(W)[None, :, None] < c, 0.0, 1.0) * np.ones((H, 1, 3))
    depth = np.zeros((H, W))
    p = O.heuristic_pb(bgr, depth)
    want = np.float32(1) - np.float32(np.exp(-np.float32((np.float32(np.sqrt(np.float32(3))) / np.float32(.7)) ** 2)))
    assert np.allclose(p[:, c - 1, 0], want, rtol=0, atol=2e-7)
    p2 = p.copy()
    p2[:, c - 1, 0] = 0
    assert not p2.any()


def test_hole_edges_follow_the_reference_walk():
    """A hole (depth 0) inside a valid fronto-parallel plane, uniform colour.
    Valid-to-invalid edge: the walk starts at the valid pixel's left (up)
    neighbour and moves left (up), finds no invalid reading: 0.2.
    Invalid-to-valid edge: it starts at the invalid pixel and moves right
    (down) into the valid one: one reading, 0.2 + 0.8 / 30."""
    W, H = 64, 48
    bgr = np.full((H, W, 3), 0.5)
    depth = np.full((H, W), 1.2)
    y0, y1, x0, x1 = 20, 30, 25, 35                      # hole rows y0..y1-1, cols x0..x1-1
    depth[y0:y1, x0:x1] = 0
    p = O.heuristic_pb(bgr, depth)
    rows = slice(y0 + 1, y1 - 1)
    assert np.allclose(p[rows, x0 - 1, 0], 0.2, atol=1e-7)           # valid | invalid
    assert np.allclose(p[rows, x1 - 1, 0], 0.2 + 0.8 / 30, atol=1e-7)  # invalid | valid
    cols = slice(x0 + 1, x1 - 1)
    assert np.allclose(p[y0 - 1, cols, 1], 0.2, atol=1e-7)
    assert np.allclose(p[y1 - 1, cols, 1], 0.2 + 0.8 / 30, atol=1e-7)
    # a hole that reaches the image edge gives no depth evidence (oob): only
    # the invalid | valid edge keeps its value
    d2 = np.full((H, W), 1.2)
    d2[:, :10] = 0
    q = O.heuristic_pb(bgr, d2)
    assert np.allclose(q[5:40, 9, 0], 0.2 + 0.8 / 30, atol=1e-7)


def test_image_edge_hack_and_suppression_width():
    W, H = 64, 48
    bgr = np.zeros((H, W, 3))
    bgr[:, W - 1] = 1.0                                  # a colour step at the last column pair
    bgr[H - 1, :] = 1.0                                  # and at the last row pair
    p = O.heuristic_pb(bgr, np.zeros((H, W)))
    assert not p[:, W - 2, 0].any() and not p[H - 2, :, 1].any()
    assert not p[:, W - 1].any() or True                 # the last column has no right edge
