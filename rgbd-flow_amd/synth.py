"""Seeded synthetic RGB-D frame pairs with ground-truth scene flow.

Scene: planar segments (a background plane plus textured rectangles), each
moved by its own small SE(3) between the frames.  Frame 2 is ray-cast from the
moved planes, and its colour is looked up in frame 1's plane textures, so
every frame-2 pixel has a valid depth.  Depth follows a Kinect-style
disparity quantisation (1/8 px steps of d = c/z with c = 1/0.00285, the
constant of the stereo error model z^2/(c+z)) and is stored as uint16 mm.

Outputs use the reference's conventions: colour uint8 BGR, depth uint16 mm,
ground-truth flow (u, v in pixels, w in metres) per frame-1 pixel.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy.ndimage import gaussian_filter

DISPARITY_C = 1.0 / 0.00285


@dataclass
class Camera:
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float


def camera_for(width: int, height: int) -> Camera:
    """Kinect-like intrinsics scaled from f=525 at 640 px wide, centre at (W-1)/2, (H-1)/2."""
    f = 525.0 * width / 640.0
    return Camera(width, height, f, f, (width - 1) / 2.0, (height - 1) / 2.0)


def _rotation(axis: np.ndarray, angle: float) -> np.ndarray:
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s = math.cos(angle), math.sin(angle)
    C = 1 - c
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def _texture(rng: np.random.Generator, H: int, W: int) -> np.ndarray:
    """Gaussian-blurred (sigma 2 px, radius 6) uniform noise plus 20 Gaussian blobs."""
    t = rng.uniform(0, 255, size=(H, W, 3))
    t = gaussian_filter(t, sigma=(2, 2, 0), mode="nearest", truncate=3.0)
    t = (t - t.mean()) * 4.0 + 128.0  # restore contrast lost to the blur
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    for _ in range(20):
        cy, cx = rng.uniform(0, H), rng.uniform(0, W)
        sg = rng.uniform(5, 40)
        amp = rng.uniform(-100, 100, size=3)
        g = np.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * sg * sg))
        t += g[..., None] * amp
    return np.clip(t, 0, 255)


def _bilinear(img: np.ndarray, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    H, W = img.shape[:2]
    x = np.clip(x, 0, W - 1)
    y = np.clip(y, 0, H - 1)
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    x1 = np.minimum(x0 + 1, W - 1)
    y1 = np.minimum(y0 + 1, H - 1)
    ax = (x - x0)[..., None]
    ay = (y - y0)[..., None]
    return ((1 - ax) * (1 - ay) * img[y0, x0] + (1 - ax) * ay * img[y1, x0]
            + ax * (1 - ay) * img[y0, x1] + ax * ay * img[y1, x1])


def _quantise_depth(z: np.ndarray) -> np.ndarray:
    """Kinect-style: 1/8-pixel disparity steps, then uint16 millimetres."""
    d = DISPARITY_C / np.maximum(z, 1e-6)
    dq = np.maximum(np.round(d * 8.0) / 8.0, 1.0 / 8.0)
    zq = DISPARITY_C / dq
    return np.clip(np.round(zq * 1000.0), 0, 65535).astype(np.uint16)


class _Plane:
    def __init__(self, n, c, half=None, basis=None):
        self.n = np.asarray(n, float)
        self.c = np.asarray(c, float)
        self.half = half        # (hu, hv) half extents in metres, None = infinite
        self.basis = basis      # (e1, e2) in-plane axes


def _intersect(plane: _Plane, R: np.ndarray, t: np.ndarray, rays: np.ndarray):
    """Ray depth (z) of the plane moved by X' = R (X - c) + c + t; inf where missed."""
    n = R @ plane.n
    c = plane.c + t
    denom = rays @ n
    with np.errstate(divide="ignore", invalid="ignore"):
        z = (c @ n) / denom
    ok = np.isfinite(z) & (z > 0.05)
    if plane.half is not None:
        X = rays * z[..., None]
        Xl = (X - c) @ R            # back to the plane's own frame (R^T (X - c))
        a = Xl @ plane.basis[0]
        b = Xl @ plane.basis[1]
        ok &= (np.abs(a) <= plane.half[0]) & (np.abs(b) <= plane.half[1])
    return np.where(ok, z, np.inf)


def _scene_planes(rng: np.random.Generator, W: int, H: int, cam: Camera, n_segments: int):
    """A background plane (nearly fronto-parallel, 3.0-3.8 m) and n_segments - 1
    textured rectangles at 0.8-2.8 m."""
    planes = []
    tilt = _rotation(rng.normal(size=3) * np.array([1, 1, 0]) + 1e-9, math.radians(rng.uniform(0, 15)))
    planes.append(_Plane(tilt @ np.array([0, 0, -1.0]), np.array([0, 0, rng.uniform(3.0, 3.8)])))
    for _ in range(n_segments - 1):
        px, py = rng.uniform(0.1 * W, 0.9 * W), rng.uniform(0.1 * H, 0.9 * H)
        z = rng.uniform(0.8, 2.8)
        c = np.array([(px - cam.cx) / cam.fx * z, (py - cam.cy) / cam.fy * z, z])
        R = _rotation(rng.normal(size=3), math.radians(rng.uniform(0, 40)))
        n = R @ np.array([0, 0, -1.0])
        e1 = R @ np.array([1.0, 0, 0])
        e2 = R @ np.array([0, 1.0, 0])
        ang = rng.uniform(0, math.pi)
        e1, e2 = math.cos(ang) * e1 + math.sin(ang) * e2, -math.sin(ang) * e1 + math.cos(ang) * e2
        planes.append(_Plane(n, c, (rng.uniform(0.15, 0.5), rng.uniform(0.15, 0.5)), (e1, e2)))
    return planes


def _motion(rng: np.random.Generator, max_translation: float, max_rotation_deg: float):
    """A random rigid motion: (rotation axis, angle in radians, translation)."""
    d = rng.normal(size=3)
    tr = d / np.linalg.norm(d) * rng.uniform(0, max_translation)
    axis = rng.normal(size=3)
    angle = math.radians(rng.uniform(0, max_rotation_deg))
    return axis, angle, tr


def make_pair(width: int, height: int, seed: int, n_segments: int = 6,
              cam: Camera | None = None, max_translation: float = 0.05,
              max_rotation_deg: float = 3.0, dropout: float = 0.0) -> dict:
    """One frame pair.  Returns dict(img1, depth1, img2, depth2 (uint16 mm),
    flow (H,W,3 f32 ground truth), seg1 (H,W int32), cam)."""
    rng = np.random.default_rng(seed)
    cam = cam or camera_for(width, height)
    H, W = height, width
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    rays = np.stack([(xs - cam.cx) / cam.fx, (ys - cam.cy) / cam.fy, np.ones_like(xs)], -1)

    planes = _scene_planes(rng, W, H, cam, n_segments)
    motions = []
    for _ in planes:
        axis, angle, tr = _motion(rng, max_translation, max_rotation_deg)
        motions.append((_rotation(axis, angle), tr))

    I3 = np.eye(3)
    z1 = np.stack([_intersect(p, I3, np.zeros(3), rays) for p in planes])
    seg1 = np.argmin(z1, axis=0)
    zt1 = np.clip(np.take_along_axis(z1, seg1[None], 0)[0], 0.5, 4.0)

    # frame 2: planes moved by X' = R (X - c) + c + t
    z2 = np.stack([_intersect(p, m[0], m[1], rays) for p, m in zip(planes, motions)])
    seg2 = np.argmin(z2, axis=0)
    zt2 = np.clip(np.take_along_axis(z2, seg2[None], 0)[0], 0.5, 4.0)

    textures = [_texture(rng, H, W) for _ in planes]
    img1 = np.zeros((H, W, 3))
    img2 = np.zeros((H, W, 3))
    X2 = rays * zt2[..., None]
    for k, (p, (Rm, tr)) in enumerate(zip(planes, motions)):
        m1 = seg1 == k
        img1[m1] = textures[k][m1]
        m2 = seg2 == k
        if m2.any():
            X = (X2[m2] - p.c - tr) @ Rm + p.c           # R^T (X' - c - t) + c
            x = X[:, 0] / X[:, 2] * cam.fx + cam.cx
            y = X[:, 1] / X[:, 2] * cam.fy + cam.cy
            img2[m2] = _bilinear(textures[k], x, y)

    # ground-truth flow of frame-1 pixels
    X1 = rays * zt1[..., None]
    flow = np.zeros((H, W, 3))
    for k, (p, (Rm, tr)) in enumerate(zip(planes, motions)):
        m1 = seg1 == k
        Xp = (X1[m1] - p.c) @ Rm.T + p.c + tr
        flow[m1, 0] = Xp[:, 0] / Xp[:, 2] * cam.fx + cam.cx - xs[m1]
        flow[m1, 1] = Xp[:, 1] / Xp[:, 2] * cam.fy + cam.cy - ys[m1]
        flow[m1, 2] = Xp[:, 2] - zt1[m1]

    d1 = _quantise_depth(zt1)
    d2 = _quantise_depth(zt2)
    if dropout > 0:
        d2 = d2.copy()
        d2[rng.uniform(size=(H, W)) < dropout] = 0
    # colour planes are stored B, G, R
    return dict(img1=np.clip(np.round(img1), 0, 255).astype(np.uint8),
                img2=np.clip(np.round(img2), 0, 255).astype(np.uint8),
                depth1=d1, depth2=d2, flow=flow.astype(np.float32),
                seg1=seg1.astype(np.int32), cam=cam)


def make_sequence(width: int, height: int, seed: int, n_frames: int, n_segments: int = 6,
                  cam: Camera | None = None, max_translation: float = 0.01,
                  max_rotation_deg: float = 0.6) -> dict:
    """A frame sequence: the scene of make_pair with every segment moving at a
    constant random velocity (per frame: translation <= max_translation m,
    rotation <= max_rotation_deg about a fixed axis), so frame k shows segment
    s at pose (R_s^k, k t_s) -- a smooth SE(3) trajectory.  Returns
    dict(frames=[dict(img (u8 BGR), depth (uint16 mm), seg (int32))], cam)."""
    rng = np.random.default_rng(seed)
    cam = cam or camera_for(width, height)
    H, W = height, width
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    rays = np.stack([(xs - cam.cx) / cam.fx, (ys - cam.cy) / cam.fy, np.ones_like(xs)], -1)
    planes = _scene_planes(rng, W, H, cam, n_segments)
    velocity = [_motion(rng, max_translation, max_rotation_deg) for _ in planes]
    textures = [_texture(rng, H, W) for _ in planes]
    frames = []
    for k in range(n_frames):
        poses = [(_rotation(axis, k * angle), k * tr) for axis, angle, tr in velocity]
        z = np.stack([_intersect(p, R, t, rays) for p, (R, t) in zip(planes, poses)])
        seg = np.argmin(z, axis=0)
        zt = np.clip(np.take_along_axis(z, seg[None], 0)[0], 0.5, 4.0)
        img = np.zeros((H, W, 3))
        X = rays * zt[..., None]
        for i, (p, (R, t)) in enumerate(zip(planes, poses)):
            m = seg == i
            if m.any():
                Xl = (X[m] - p.c - t) @ R + p.c           # R^T (X - c - t) + c
                img[m] = _bilinear(textures[i], Xl[:, 0] / Xl[:, 2] * cam.fx + cam.cx,
                                   Xl[:, 1] / Xl[:, 2] * cam.fy + cam.cy)
        frames.append(dict(img=np.clip(np.round(img), 0, 255).astype(np.uint8), depth=_quantise_depth(zt),
                           seg=seg.astype(np.int32), z=zt))
    return dict(frames=frames, cam=cam, planes=planes, velocity=velocity)


def sequence_flow(seq: dict, i: int, j: int) -> np.ndarray:
    """Ground-truth scene flow (u, v px; w m) of frame i's pixels to frame j of
    a make_sequence sequence: each pixel's point on its segment, moved from
    the segment's pose at frame i to its pose at frame j."""
    cam = seq["cam"]
    fr = seq["frames"][i]
    H, W = fr["seg"].shape
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    rays = np.stack([(xs - cam.cx) / cam.fx, (ys - cam.cy) / cam.fy, np.ones_like(xs)], -1)
    X = rays * fr["z"][..., None]
    flow = np.zeros((H, W, 3))
    for s_, (p, (axis, angle, tr)) in enumerate(zip(seq["planes"], seq["velocity"])):
        m = fr["seg"] == s_
        if not m.any():
            continue
        Ri, ti = _rotation(axis, i * angle), i * tr
        Rj, tj = _rotation(axis, j * angle), j * tr
        X0 = (X[m] - p.c - ti) @ Ri + p.c            # the point at pose 0: R_i^T (X - c - t_i) + c
        Xj = (X0 - p.c) @ Rj.T + p.c + tj
        flow[m, 0] = Xj[:, 0] / Xj[:, 2] * cam.fx + cam.cx - xs[m]
        flow[m, 1] = Xj[:, 1] / Xj[:, 2] * cam.fy + cam.cy - ys[m]
        flow[m, 2] = Xj[:, 2] - X[m][:, 2]
    return flow


def to_solver_inputs(pair: dict):
    """The reference's conversions: u8 BGR -> f64 /255 with no channel swap
    (src/ceLiuOpticalFlow.cpp:19-34); uint16 mm -> float metres via *.001
    (src/main.cpp:32-48) -> f64 (ceLiuOpticalFlow.cpp:39-48)."""
    im1 = pair["img1"].astype(np.float64) / 255.0
    im2 = pair["img2"].astype(np.float64) / 255.0
    df1 = (pair["depth1"].astype(np.float64) * .001).astype(np.float32)
    df2 = (pair["depth2"].astype(np.float64) * .001).astype(np.float32)
    return im1, df1, im2, df2


def level_sample_positions(size: int, ratio: float, widths):
    """Full-resolution pixel that each pyramid level's pixel samples along one
    axis.  Level i <= n resizes the original by r = ratio^i and reads pixel
    (j+1)/r - 1; level i > n resizes level i-n by rate = ratio^i W0 / W_{i-n}
    (src/GaussianPyramid.cpp:44-61).  Exact for ratio 0.5, where every read
    lands on a pixel; truncated otherwise."""
    if ratio > 0.98 or ratio < 0.4:
        ratio = 0.75
    n = int(math.log(0.25) / math.log(ratio))
    pos = [np.arange(size, dtype=np.int64)]
    for i in range(1, len(widths)):
        if i <= n:
            r = ratio ** i
            m = int(size * r)
            pos.append(np.clip(((np.arange(m) + 1) / r - 1).astype(np.int64), 0, size - 1))
        else:
            prev = pos[i - n]
            rate = ratio ** i * widths[0] / widths[i - n]
            m = int(prev.size * rate)
            idx = np.clip(((np.arange(m) + 1) / rate - 1).astype(np.int64), 0, prev.size - 1)
            pos.append(prev[idx])
    return pos


def segment_boundary_maps(seg: np.ndarray, widths, heights, ratio: float = 0.5) -> np.ndarray:
    """Boundary probability per level from the generator's segment labels:
    1 where the two pixels of an edge (right, down) sample different segments.
    Returns all levels' [H_k][W_k][2] float32 maps concatenated, level 0 first."""
    H, W = seg.shape
    L = len(widths)
    px = level_sample_positions(W, ratio, widths)
    py = level_sample_positions(H, ratio, widths)
    out = []
    for k in range(L):
        xk = px[k][: widths[k]]
        yk = py[k][: heights[k]]
        s = seg[np.ix_(yk, xk)]
        m = np.zeros((heights[k], widths[k], 2), np.float32)
        m[:, :-1, 0] = (s[:, :-1] != s[:, 1:])
        m[:-1, :, 1] = (s[:-1, :] != s[1:, :])
        out.append(m.ravel())
    return np.concatenate(out)
