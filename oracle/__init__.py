"""ctypes binding of the oracle: the CPU restatement of rgbd-flow's scene-flow path.

TEST INFRASTRUCTURE ONLY -- the checker.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module.  The product library
(librgbdflow.so) never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

PREC_REF, PREC_F64 = 0, 1
REG_ISOTROPIC, REG_LETOUZEY, REG_HPB = 0, 1, 2
IMG_F64, IMG_F16, IMG_BF16 = 0, 1, 2
REC_F32, REC_F16, REC_BF16 = 0, 1, 2


class Camera(C.Structure):
    _fields_ = [("xRes", C.c_uint), ("yRes", C.c_uint), ("centerX", C.c_float), ("centerY", C.c_float),
                ("focalLength", C.c_float), ("stereoBaseline", C.c_float), ("fovX", C.c_float),
                ("fovY", C.c_float)]


class Params(C.Structure):
    _fields_ = [("regularization_type", C.c_int), ("color_data_weight", C.c_double),
                ("depth_vs_color_data_weight", C.c_double), ("alpha", C.c_double), ("gamma", C.c_double),
                ("pyramid_ratio", C.c_double), ("min_width", C.c_int), ("n_outer", C.c_int),
                ("n_inner", C.c_int), ("n_sor", C.c_int), ("cam", Camera), ("precision", C.c_int),
                ("disable_inner_exit", C.c_int), ("disable_outer_exit", C.c_int), ("omega", C.c_float),
                ("residual_exit", C.c_int), ("residual_tol", C.c_double), ("pixel_gs", C.c_int),
                ("image_storage", C.c_int), ("record_storage", C.c_int), ("cg_levels", C.c_int),
                ("cg_sweeps", C.c_int), ("cg_at", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("n_levels", C.c_int), ("width", C.c_int * 16), ("height", C.c_int * 16),
                ("outer_exec", C.c_int * 16), ("sweeps_exec", C.c_int * 16), ("last_stat", C.c_double * 16),
                ("residual_checks", C.c_int * 16), ("residual", C.c_double * 16), ("energy", C.c_double * 16)]


class LevelIn(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("focal_level", C.c_float),
                ("I1", C.c_void_p), ("I2", C.c_void_p), ("warpI2", C.c_void_p),
                ("D1", C.c_void_p), ("D2", C.c_void_p), ("warpD2", C.c_void_p),
                ("delta", C.c_void_p), ("s", C.c_void_p), ("m", C.c_void_p), ("r", C.c_void_p)]


class LevelStats(C.Structure):
    _fields_ = [("outer_exec", C.c_int), ("sweeps_exec", C.c_int), ("last_stat", C.c_double),
                ("last_maxdiff", C.c_float), ("residual_checks", C.c_int), ("last_residual", C.c_double),
                ("last_energy", C.c_double)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_pyramid_levels.restype = C.c_int
        _lib.orc_level_solve.restype = C.c_int
        _lib.orc_multiscale.restype = C.c_int
        _lib.orc_with_init.restype = C.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def make_params(W, H, f=None, cx=None, cy=None, alpha=0.02, gamma=1e-5, n_outer=15, n_inner=1, n_sor=50,
                ratio=0.5, min_width=None, reg=REG_HPB, precision=PREC_REF, cw=1.0, dvc=1.0,
                disable_inner_exit=False, disable_outer_exit=False, omega=0.7, residual_exit=False,
                residual_tol=1e-4, pixel_gs=False, image_storage=IMG_F64, record_storage=0, cg_levels=0, cg_sweeps=2,
                cg_at=9) -> Params:
    p = Params()
    p.regularization_type = reg
    p.color_data_weight = cw
    p.depth_vs_color_data_weight = dvc
    p.alpha = alpha
    p.gamma = gamma
    p.pyramid_ratio = ratio
    p.min_width = min_width if min_width is not None else W // 16
    p.n_outer, p.n_inner, p.n_sor = n_outer, n_inner, n_sor
    p.cam.xRes, p.cam.yRes = W, H
    p.cam.focalLength = f if f is not None else 525.0 * W / 640.0
    p.cam.centerX = cx if cx is not None else (W - 1) / 2.0
    p.cam.centerY = cy if cy is not None else (H - 1) / 2.0
    p.precision = precision
    p.disable_inner_exit = int(disable_inner_exit)
    p.disable_outer_exit = int(disable_outer_exit)
    p.omega = omega
    p.residual_exit = int(residual_exit)
    p.residual_tol = residual_tol
    p.pixel_gs = int(pixel_gs)
    p.image_storage = image_storage
    p.record_storage = record_storage
    p.cg_levels, p.cg_sweeps, p.cg_at = cg_levels, cg_sweeps, cg_at
    return p


def img_round(x, storage):
    """x rounded as a feature image of that storage holds it (f32, then FP16 / BF16)."""
    f = lib().orc_img_round
    f.restype = C.c_double
    f.argtypes = [C.c_double, C.c_int]
    return np.array([f(float(v), storage) for v in np.ravel(x)]).reshape(np.shape(x))


def record_round(c, storage):
    """Rows of 13 f32 system coefficients rounded as that record storage holds them."""
    a = np.array(c, dtype=np.float32).reshape(-1, 13).copy()
    f = lib().orc_record_round
    f.argtypes = [C.c_void_p, C.c_int]
    for row in a:
        f(row.ctypes.data, storage)
    return a


# ---------------------------------------------------------------- pyramids
def pyramid_levels(W, H, ratio, min_width):
    ws = (C.c_int * 16)()
    hs = (C.c_int * 16)()
    n = lib().orc_pyramid_levels(W, H, C.c_double(ratio), min_width, ws, hs)
    return [(ws[i], hs[i]) for i in range(n)]


def gaussian_smooth(img, sigma, fsize):
    img = _f64(img)
    H, W = img.shape[:2]
    Cn = 1 if img.ndim == 2 else img.shape[2]
    out = np.empty_like(img)
    lib().orc_gaussian_smooth(_p(img), W, H, Cn, C.c_double(sigma), fsize, _p(out))
    return out


def resize_ratio(img, ratio):
    img = _f64(img)
    H, W = img.shape[:2]
    Cn = 1 if img.ndim == 2 else img.shape[2]
    DW, DH = int(W * ratio), int(H * ratio)
    out = np.empty((DH, DW) + img.shape[2:], np.float64)
    lib().orc_resize_ratio(_p(img), W, H, Cn, C.c_double(ratio), _p(out))
    return out


def resize_dims(img, Wd, Hd):
    img = _f64(img)
    H, W = img.shape[:2]
    Cn = 1 if img.ndim == 2 else img.shape[2]
    out = np.empty((Hd, Wd) + img.shape[2:], np.float64)
    lib().orc_resize_dims(_p(img), W, H, Cn, Wd, Hd, _p(out))
    return out


def _split_levels(flat, levels, Cn):
    out, off = [], 0
    for (w, h) in levels:
        n = w * h * Cn
        a = flat[off:off + n]
        out.append(a.reshape((h, w) + ((Cn,) if Cn > 1 else ())))
        off += n
    return out


def gaussian_pyramid(img, ratio, min_width):
    img = _f64(img)
    H, W = img.shape[:2]
    Cn = 1 if img.ndim == 2 else img.shape[2]
    levels = pyramid_levels(W, H, ratio, min_width)
    flat = np.empty(sum(w * h for w, h in levels) * Cn, np.float64)
    lib().orc_gaussian_pyramid(_p(img), W, H, Cn, C.c_double(ratio), min_width, _p(flat))
    return _split_levels(flat, levels, Cn)


def depth_map_pyramid(depth, ratio, min_width):
    depth = _f64(depth)
    H, W = depth.shape
    levels = pyramid_levels(W, H, ratio, min_width)
    sizes = [(int(W * ratio ** m), int(H * ratio ** m)) for m in range(len(levels))]
    flat = np.empty(sum(w * h for w, h in sizes), np.float64)
    lib().orc_depth_map_pyramid(_p(depth), W, H, C.c_double(ratio), min_width, _p(flat))
    return _split_levels(flat, sizes, 1)


# ------------------------------------------------------- features and weights
def im2feature(bgr):
    bgr = _f64(bgr)
    H, W = bgr.shape[:2]
    out = np.empty((H, W, 5), np.float64)
    lib().orc_im2feature(_p(bgr), W, H, _p(out))
    return out


def warp_cpu(I1, D1, I2, D2, u, v, w):
    I1, D1, I2, D2, u, v, w = map(_f64, (I1, D1, I2, D2, u, v, w))
    H, W = D1.shape
    Cn = I1.shape[2]
    wI = np.empty_like(I1)
    wD = np.empty_like(D1)
    lib().orc_warp_cpu(_p(I1), _p(D1), _p(I2), _p(D2), _p(u), _p(v), _p(w), W, H, Cn, _p(wI), _p(wD))
    return wI, wD


def depth_stdevs(depth_f32, max_dist=4):
    d = _f32(depth_f32)
    H, W = d.shape
    out = np.empty_like(d)
    lib().orc_depth_stdevs_pointwise(_p(d), W, H, max_dist, _p(out))
    return out


def l1_distance_to_boundary(depth_f32, cap=4):
    d = _f32(depth_f32)
    H, W = d.shape
    out = np.empty_like(d)
    lib().orc_l1_distance_to_boundary(_p(d), W, H, cap, _p(out))
    return out


def depth_data_weights(sigma, dvc=1.0):
    s = _f32(sigma)
    H, W = s.shape
    out = np.empty((H, W), np.float64)
    lib().orc_depth_data_weights(_p(s), W, H, C.c_double(dvc), _p(out))
    return out


def depth_mag_weights(depth_level, W_full, H_full, f_full):
    d = _f64(depth_level)
    H, W = d.shape
    out = np.empty((H, W), np.float64)
    lib().orc_depth_mag_weights(_p(d), W, H, W_full, H_full, C.c_float(f_full), _p(out))
    return out


def reg_weights(hpb, depth, reg=REG_HPB):
    d = _f64(depth)
    H, W = d.shape
    h = _f32(hpb) if hpb is not None else np.zeros((H, W, 2), np.float32)
    out = np.empty((H, W, 2), np.float64)
    lib().orc_reg_weights(_p(h), _p(d), W, H, reg, _p(out))
    return out


def median_fill(depth, halfwidth=1):
    d = _f64(depth).copy()
    H, W = d.shape
    lib().orc_median_fill(_p(d), W, H, halfwidth)
    return d


def colour_fill_rounds(depth, bgr, max_region=60):
    """The device's rounds variant of the colour-guided small-hole fill:
    (filled depth, pixels of small regions left invalid).  bgr: (H, W, 3)."""
    d = _f64(depth).copy()
    H, W = d.shape
    planes = np.ascontiguousarray(np.moveaxis(_f64(bgr), -1, 0))
    left = lib().orc_colour_fill_rounds(_p(d), _p(planes), W, H, max_region)
    return d, left


def fill_exp(x):
    f = lib().orc_fill_exp
    f.restype = C.c_double
    f.argtypes = [C.c_double]
    return np.array([f(float(t)) for t in np.ravel(x)])


def median_fill_rounds(depth):
    """The device's parallel-rounds hole fill: (filled depth, pixels left invalid)."""
    d = _f64(depth).copy()
    H, W = d.shape
    left = lib().orc_median_fill_rounds(_p(d), W, H)
    return d, left


# ------------------------------------------------------------------ solvers
class LevelInputs:
    """Arrays of one pyramid level, in the reference's DImage layouts."""

    def __init__(self, I1, I2, warpI2, D1, D2, warpD2, delta, s, m, r, focal_level):
        self.I1, self.I2, self.warpI2 = _f64(I1), _f64(I2), _f64(warpI2)
        self.D1, self.D2, self.warpD2 = _f64(D1), _f64(D2), _f64(warpD2)
        self.delta, self.s, self.m, self.r = _f64(delta), _f64(s), _f64(m), _f64(r)
        self.focal_level = focal_level
        self.H, self.W = self.D1.shape

    def struct(self):
        li = LevelIn()
        li.W, li.H, li.focal_level = self.W, self.H, self.focal_level
        for name in ("I1", "I2", "warpI2", "D1", "D2", "warpD2", "delta", "s", "m", "r"):
            setattr(li, name, getattr(self, name).ctypes.data)
        return li


def level_solve(li: LevelInputs, params: Params, allow_early, u, v, w):
    u, v, w = _f64(u).copy(), _f64(v).copy(), _f64(w).copy()
    st = LevelStats()
    s = li.struct()
    lib().orc_level_solve(C.byref(s), C.byref(params), int(allow_early), _p(u), _p(v), _p(w), C.byref(st))
    return u, v, w, st


def warp_gpu_style(I1, D1, I2, D2, u, v, w, precision=PREC_REF):
    I1, D1, I2, D2, u, v, w = map(_f64, (I1, D1, I2, D2, u, v, w))
    H, W = D1.shape
    wI = np.empty_like(I1)
    wD = np.empty_like(D1)
    lib().orc_warp_gpu_style(_p(I1), _p(D1), _p(I2), _p(D2), _p(u), _p(v), _p(w), W, H, precision, _p(wI), _p(wD))
    return wI, wD


def warp_gpu_style_f32(I1, D1, I2, D2, u, v, w, image_storage):
    I1, D1, I2, D2, u, v, w = map(_f64, (I1, D1, I2, D2, u, v, w))
    H, W = D1.shape
    wI = np.empty_like(I1)
    wD = np.empty_like(D1)
    lib().orc_warp_gpu_style_f32(_p(I1), _p(D1), _p(I2), _p(D2), _p(u), _p(v), _p(w), W, H, image_storage, _p(wI),
                                 _p(wD))
    return wI, wD


def level_weights(li: LevelInputs, params: Params, u, v, w):
    u, v, w = _f64(u), _f64(v), _f64(w)
    phi = np.empty((li.H, li.W)); psi = np.empty((li.H, li.W, 5)); pi = np.empty((li.H, li.W))
    s = li.struct()
    lib().orc_level_weights(C.byref(s), C.byref(params), _p(u), _p(v), _p(w), _p(phi), _p(psi), _p(pi))
    return phi, psi, pi


def level_system(li: LevelInputs, params: Params, u, v, w):
    u, v, w = _f64(u), _f64(v), _f64(w)
    out = np.empty((13, li.H, li.W), np.float64)
    s = li.struct()
    lib().orc_level_system(C.byref(s), C.byref(params), _p(u), _p(v), _p(w), _p(out))
    return out


def relax(coef13, n_iter, x, precision=PREC_REF):
    c = _f64(coef13)
    _, H, W = c.shape
    xx = _f64(x).copy()
    lib().orc_relax(_p(c), W, H, precision, n_iter, _p(xx))
    return xx


def heuristic_pb(bgr, depth):
    """computeHeuristicPb of one level: (H, W, 3) BGR in [0,1], (H, W) metres
    -> (H, W, 2) f32 boundary probabilities (right, down)."""
    b = _f64(bgr)
    d = _f64(depth)
    H, W = d.shape
    out = np.empty((H, W, 2), np.float32)
    lib().orc_heuristic_pb(_p(b), _p(d), W, H, _p(out))
    return out


def hpb_normals(depth):
    d = _f64(depth)
    H, W = d.shape
    n = np.empty((H, W, 3), np.float32)
    v = np.empty((H, W), np.uint32)
    lib().orc_hpb_normals(_p(d), W, H, _p(n), _p(v))
    return n, v


def hpb_exp(x):
    f = lib().orc_hpb_exp
    f.restype = C.c_float
    f.argtypes = [C.c_float]
    return np.array([f(float(t)) for t in np.asarray(x, np.float32).ravel()], np.float32)


def hpb_smallest_eigvec(a):
    """The normals' eigenvector of the smallest eigenvalue of a symmetric PSD 3x3 (f64)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = np.zeros(3)
    lib().orc_hpb_smallest_eigvec(_p(a), _p(n))
    return n


def cg_operator(coef13, precision=PREC_REF):
    """The coarse-grid correction's coarse system of a compact system (rhs zero)."""
    c = _f64(coef13)
    _, H, W = c.shape
    out = np.empty((13, (H + 1) // 2, (W + 1) // 2))
    lib().orc_cg_operator(_p(c), W, H, precision, _p(out))
    return out


def cg_restrict(coef13, x, precision=PREC_REF):
    """The coarse rhs: the residual at x ([3][H][W]) summed over 2x2 aggregates."""
    c, xx = _f64(coef13), _f64(x)
    _, H, W = c.shape
    out = np.empty((3, (H + 1) // 2, (W + 1) // 2))
    lib().orc_cg_restrict(_p(c), W, H, precision, _p(xx), _p(out))
    return out


def cg_prolong(x, xc, precision=PREC_REF):
    """x + P xc."""
    xx, cc = np.array(x, np.float64, copy=True), _f64(xc)
    _, H, W = xx.shape
    lib().orc_cg_prolong(_p(xx), W, H, precision, _p(cc))
    return xx


def relax_cg(coef13, n_iter, x, params: Params):
    """n_iter relaxation iterations with params' coarse-grid correction; returns
    (x, per-iteration max change, per-iteration f64 relative residual)."""
    c = _f64(coef13)
    _, H, W = c.shape
    xx = np.array(x, np.float64, copy=True)
    md = np.empty(n_iter)
    rs = np.empty(n_iter)
    lib().orc_relax_cg(_p(c), W, H, C.byref(params), n_iter, _p(xx), _p(md), _p(rs))
    return xx, md, rs


def system_residual(coef13, x):
    """(relative residual ||r - M x|| / ||r||, energy r'x - 1/2 x'Mx) in f64."""
    c = _f64(coef13)
    _, H, W = c.shape
    xx = _f64(x)
    out = np.zeros(2)
    lib().orc_system_residual(_p(c), W, H, _p(xx), _p(out))
    return float(out[0]), float(out[1])


def relax_residual(coef13, max_iter, tol, x, omega=0.7, precision=PREC_REF, pixel_gs=False):
    """Relax until the residual test passes at a check iteration; returns
    (x, iterations, residual, energy)."""
    c = _f64(coef13)
    _, H, W = c.shape
    xx = _f64(x).copy()
    it, res, en = C.c_int(), C.c_double(), C.c_double()
    lib().orc_relax_residual(_p(c), W, H, precision, C.c_float(omega), int(pixel_gs), int(max_iter),
                             C.c_double(tol), _p(xx), C.byref(it), C.byref(res), C.byref(en))
    return xx, it.value, res.value, en.value


def multiscale(im1, dm1, im2, dm2, sigma1, hpb, params: Params):
    im1, dm1, im2, dm2 = _f64(im1), _f64(dm1), _f64(im2), _f64(dm2)
    s1 = _f32(sigma1)
    H, W = dm1.shape
    h = _f32(hpb) if hpb is not None else None
    u = np.empty((H, W)); v = np.empty((H, W)); w = np.empty((H, W))
    st = Stats()
    rc = lib().orc_multiscale(_p(im1), _p(dm1), _p(im2), _p(dm2), _p(s1), _p(h) if h is not None else None,
                              W, H, C.byref(params), _p(u), _p(v), _p(w), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_multiscale failed ({rc})")
    return u, v, w, st


def with_init(im1, dm1, im2, dm2, sigma1, hpb0, est, params: Params):
    im1, dm1, im2, dm2, est = _f64(im1), _f64(dm1), _f64(im2), _f64(dm2), _f64(est)
    s1 = _f32(sigma1)
    H, W = dm1.shape
    h = _f32(hpb0) if hpb0 is not None else None
    u = np.empty((H, W)); v = np.empty((H, W)); w = np.empty((H, W))
    st = Stats()
    rc = lib().orc_with_init(_p(im1), _p(dm1), _p(im2), _p(dm2), _p(s1), _p(h) if h is not None else None,
                             _p(est), W, H, C.byref(params), _p(u), _p(v), _p(w), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"orc_with_init failed ({rc})")
    return u, v, w, st


def set_threads(n):
    """Threads of the level solve's row partitions (1 = scalar, the default);
    results do not depend on the count."""
    L = lib()
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_set_threads(int(n))


def threads():
    L = lib()
    L.orc_threads.restype = C.c_int
    return int(L.orc_threads())


def check_reciprocal_division(n, seed, mode, recipe=0):
    """Samples n (a, b) in the reciprocal division's domain (divcheck.cpp) and
    returns (mismatches against IEEE a / b, first mismatching (a, b))."""
    L = lib()
    f = L.orc_check_reciprocal_division
    f.restype = C.c_uint64
    f.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    a, b = C.c_float(), C.c_float()
    bad = f(n, seed, mode, recipe, C.byref(a), C.byref(b))
    return int(bad), (a.value, b.value)


def check_div_const(n, seed, d, recipe=0):
    """Samples n f64 x in the domain of the kernels' division by the integer
    constant d (divcheck.cpp) and returns (mismatches against IEEE x / d,
    first x)."""
    L = lib()
    f = L.orc_check_div_const
    f.restype = C.c_uint64
    f.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_double)]
    x = C.c_double()
    bad = f(n, seed, d, recipe, C.byref(x))
    return int(bad), x.value


def check_div_const_f32(d, recipe=0):
    """Every f32 x of the binade [1, 2) (both signs) through the f32 division
    by the constant d (divcheck.cpp); returns (mismatches against IEEE x / d,
    first x)."""
    L = lib()
    f = L.orc_check_div_const_f32
    f.restype = C.c_uint64
    f.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_float)]
    x = C.c_float()
    bad = f(d, recipe, C.byref(x))
    return int(bad), x.value


def check_recip_short(bits, recipe=0):
    """Every significand of `bits` significant bits of the binade [1, 2)
    (both signs) through the kernels' refined reciprocal of a 16-bit record's
    diagonal, from each y0 within one ulp of 1/c (divcheck.cpp); returns
    (mismatches against IEEE 1.0f / c, first c)."""
    L = lib()
    f = L.orc_check_recip_short
    f.restype = C.c_uint64
    f.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_float)]
    x = C.c_float()
    bad = f(bits, recipe, C.byref(x))
    return int(bad), x.value
