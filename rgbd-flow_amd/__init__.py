"""Python binding of librgbdflow's C ABI (include/rgbdflow/rgbdflow.h).

A harness for tests and benchmarks: every call goes through the C ABI into
the gfx950 HIP kernels.  There is no Python or CPU fallback -- if the native
library is missing, or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librgbdflow.so")
LIB_CXX_PATH = os.path.join(HERE, "librgbdflow_cxx.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_BUDGET, ERR_STATE, ERR_UNSUPPORTED = -1, -2, -3, -4, -5
REG_ISOTROPIC, REG_LETOUZEY, REG_HPB = 0, 1, 2
VARIANT_FULL = 2
RELAX_PLAIN, RELAX_TILED = 0, 1
STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2
EXIT_NONE, EXIT_REFERENCE, EXIT_RESIDUAL = 0, 1, 2
PIXEL_JACOBI, PIXEL_GS, PIXEL_BLOCK = 0, 1, 2
IMAGES_F64, IMAGES_F16, IMAGES_BF16 = 0, 1, 2

# Every symbol include/rgbdflow/rgbdflow.h declares.
EXPORTS = [
    "rgbdflow_default_params", "rgbdflow_create", "rgbdflow_destroy", "rgbdflow_last_error",
    "rgbdflow_bytes_in_use", "rgbdflow_synchronize", "rgbdflow_set_relax_timing", "rgbdflow_pyramid_levels",
    "rgbdflow_solve_level", "rgbdflow_depth_stdevs_pointwise", "rgbdflow_fill_depth_holes", "rgbdflow_fill_small_holes_colour",
    "rgbdflow_pair_load", "rgbdflow_pair_run", "rgbdflow_pair_read", "rgbdflow_multiscale",
    "rgbdflow_with_init", "rgbdflow_gaussian_pyramid", "rgbdflow_depth_map_pyramid", "rgbdflow_im2feature",
    "rgbdflow_warp_by_3d_flow", "rgbdflow_upsample_flow", "rgbdflow_level_system", "rgbdflow_relax",
    "rgbdflow_relax_residual", "rgbdflow_heuristic_pb", "rgbdflow_autotune_relax", "rgbdflow_batch_load",
    "rgbdflow_batch_run", "rgbdflow_batch_read", "rgbdflow_set_relax_schedule", "rgbdflow_time_relax_schedule",
    "rgbdflow_server_create", "rgbdflow_server_destroy", "rgbdflow_server_last_error", "rgbdflow_stream_open",
    "rgbdflow_stream_submit", "rgbdflow_stream_poll", "rgbdflow_stream_close",
    "rgbdflow_server_set_relax_schedule", "rgbdflow_fill_depth_holes_noalloc", "rgbdflow_relax_shape",
]


class Camera(C.Structure):
    _fields_ = [("xRes", C.c_uint), ("yRes", C.c_uint), ("centerX", C.c_float), ("centerY", C.c_float),
                ("focalLength", C.c_float), ("stereoBaseline", C.c_float), ("fovX", C.c_float),
                ("fovY", C.c_float)]


class Params(C.Structure):
    _fields_ = [("variant", C.c_int), ("regularization_type", C.c_int), ("cam", Camera),
                ("color_data_weight", C.c_double), ("depth_vs_color_data_weight", C.c_double),
                ("alpha", C.c_double), ("gamma", C.c_double), ("pyramid_ratio", C.c_double),
                ("min_width", C.c_int), ("n_outer", C.c_int), ("n_inner", C.c_int), ("n_sor", C.c_int),
                ("relax_variant", C.c_int), ("inner_exit", C.c_int), ("outer_exit", C.c_int),
                ("tile_iters", C.c_int), ("tile_w", C.c_int), ("tile_h", C.c_int), ("tile_config", C.c_int),
                ("graph", C.c_int), ("storage", C.c_int), ("omega", C.c_float), ("residual_tol", C.c_float),
                ("pixel_update", C.c_int), ("image_storage", C.c_int), ("coarse_levels", C.c_int),
                ("coarse_sweeps", C.c_int), ("coarse_at", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("n_levels", C.c_int), ("width", C.c_int * 16), ("height", C.c_int * 16),
                ("outer_exec", C.c_int * 16), ("sweeps_exec", C.c_int * 16), ("last_stat", C.c_double * 16),
                ("device_ms", C.c_double), ("relax_ms", C.c_double), ("relax_launches", C.c_int),
                ("residual_checks", C.c_int * 16), ("residual", C.c_double * 16), ("energy", C.c_double * 16),
                ("stat_margin", C.c_double * 16)]

    def levels(self):
        return [dict(W=self.width[k], H=self.height[k], outer=self.outer_exec[k], sweeps=self.sweeps_exec[k],
                     stat=self.last_stat[k], residual_checks=self.residual_checks[k], residual=self.residual[k],
                     energy=self.energy[k], stat_margin=self.stat_margin[k]) for k in range(self.n_levels)]


class PairInput(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("im1", C.c_void_p), ("dm1", C.c_void_p), ("im2", C.c_void_p),
                ("dm2", C.c_void_p), ("sigma1", C.c_void_p), ("hpb", C.c_void_p), ("cam", Camera),
                ("min_width", C.c_int), ("reserved", C.c_int), ("est", C.c_void_p), ("hpb_dm1", C.c_void_p)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("max_batch", C.c_int), ("lanes", C.c_int), ("warm_start", C.c_int),
                ("hbm_budget", C.c_ulonglong), ("warm_n_outer", C.c_int), ("fill_holes", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("has_flow", C.c_int), ("status", C.c_int), ("batch_pairs", C.c_int), ("stream", C.c_int),
                ("queue_ms", C.c_double), ("batch_ms", C.c_double), ("latency_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TileChoice(C.Structure):
    _fields_ = [("tile_w", C.c_int), ("tile_h", C.c_int), ("tile_iters", C.c_int), ("tile_config", C.c_int),
                ("threads", C.c_int), ("pairs_per_thread", C.c_int), ("single", C.c_int), ("candidates", C.c_int),
                ("us_per_iter", C.c_double), ("streamed", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def relax_shapes():
    """Workgroup shapes of the tiled relaxation (tile_config index order):
    dicts of threads, pairs_per_thread and streamed (the most iterations per
    launch of a streamed shape, else 0)."""
    out = []
    while True:
        t, p, s = C.c_int(), C.c_int(), C.c_int()
        if lib().rgbdflow_relax_shape(len(out), C.byref(t), C.byref(p), C.byref(s)) != 0:
            return out
        out.append(dict(threads=t.value, pairs_per_thread=p.value, streamed=s.value))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"librgbdflow is not built ({LIB_PATH}); run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _lib.rgbdflow_last_error.restype = C.c_char_p
        _lib.rgbdflow_bytes_in_use.restype = C.c_ulonglong
        _lib.rgbdflow_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_ulonglong]
        _lib.rgbdflow_server_last_error.restype = C.c_char_p
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class RgbdFlowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rgbdflow error {code}: {msg}")
        self.code = code


def make_params(W, H, f=None, cx=None, cy=None, alpha=0.02, gamma=1e-5, n_outer=15, n_inner=1, n_sor=50,
                ratio=0.5, min_width=None, reg=REG_HPB, cw=1.0, dvc=1.0, inner_exit=True, outer_exit=True,
                relax_variant=RELAX_PLAIN, tile_iters=0, tile_w=0, tile_h=0, tile_config=-1, graph=False,
                storage=STORE_F32, omega=0.7, residual_tol=1e-4, pixel_update=PIXEL_JACOBI,
                image_storage=IMAGES_F64, coarse_levels=0, coarse_sweeps=2, coarse_at=9) -> Params:
    """inner_exit: True / EXIT_REFERENCE (the reference's max-change test),
    False / EXIT_NONE, or EXIT_RESIDUAL (FP64 relative residual <= residual_tol)."""
    p = Params()
    lib().rgbdflow_default_params(C.byref(p))
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
    p.inner_exit = int(inner_exit)
    p.outer_exit = int(outer_exit)
    p.relax_variant = relax_variant
    p.tile_iters = tile_iters
    p.tile_w, p.tile_h, p.tile_config = tile_w, tile_h, tile_config
    p.graph = int(graph)
    p.storage = storage
    p.omega = omega
    p.residual_tol = residual_tol
    p.pixel_update = pixel_update
    p.image_storage = image_storage
    p.coarse_levels, p.coarse_sweeps, p.coarse_at = coarse_levels, coarse_sweeps, coarse_at
    return p


def pyramid_levels(W, H, ratio, min_width):
    ws = (C.c_int * 16)()
    hs = (C.c_int * 16)()
    n = lib().rgbdflow_pyramid_levels(W, H, C.c_double(ratio), min_width, ws, hs)
    if n < 0:
        raise RgbdFlowError(n, "bad pyramid arguments")
    return [(ws[i], hs[i]) for i in range(n)]


class Context:
    """One librgbdflow context (HIP stream + pooled HBM buffers) on one device."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        rc = lib().rgbdflow_create(C.byref(h), device, hbm_budget)
        if rc != OK:
            raise RgbdFlowError(rc, "rgbdflow_create failed (no HIP device?)")
        self.h = h
        self._batch_shapes = []
        self._pair_shape = None

    def close(self):
        if self.h:
            lib().rgbdflow_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise RgbdFlowError(rc, lib().rgbdflow_last_error(self.h).decode())

    def bytes_in_use(self):
        return lib().rgbdflow_bytes_in_use(self.h)

    def synchronize(self):
        self._check(lib().rgbdflow_synchronize(self.h))

    def autotune_relax(self, W, H, batch=1, params: Params | None = None):
        """Time the tiled relaxation's schedules on `batch` W x H levels (systems
        stored as params.storage); the fastest becomes this context's default
        for that shape."""
        ch = TileChoice()
        self._check(lib().rgbdflow_autotune_relax(self.h, W, H, int(batch),
                                                  C.byref(params) if params is not None else None, C.byref(ch)))
        return ch.as_dict()

    def set_relax_schedule(self, W, H, choice: dict):
        """Install a schedule (a dict as autotune_relax returns) for W x H levels."""
        ch = TileChoice()
        for f, _ in TileChoice._fields_:
            if f in choice:
                setattr(ch, f, choice[f])
        self._check(lib().rgbdflow_set_relax_schedule(self.h, W, H, C.byref(ch)))

    def time_relax_schedule(self, W, H, choice: dict, batch=1, params: Params | None = None):
        """Device microseconds per red-black iteration of a schedule."""
        ch = TileChoice()
        for f, _ in TileChoice._fields_:
            if f in choice:
                setattr(ch, f, choice[f])
        us = C.c_double()
        self._check(lib().rgbdflow_time_relax_schedule(self.h, W, H, int(batch),
                                                       C.byref(params) if params is not None else None,
                                                       C.byref(ch), C.byref(us)))
        return us.value

    def set_relax_timing(self, on: bool):
        self._check(lib().rgbdflow_set_relax_timing(self.h, int(on)))

    # ---------------------------------------------------------------- stages
    def gaussian_pyramid(self, img, ratio, min_width):
        img = _f64(img)
        H, W = img.shape[:2]
        Cn = 1 if img.ndim == 2 else img.shape[2]
        levels = pyramid_levels(W, H, ratio, min_width)
        flat = np.empty(sum(w * h for w, h in levels) * Cn, np.float64)
        self._check(lib().rgbdflow_gaussian_pyramid(self.h, _p(img), W, H, Cn, C.c_double(ratio), min_width,
                                                    _p(flat)))
        out, off = [], 0
        for (w, h) in levels:
            n = w * h * Cn
            out.append(flat[off:off + n].reshape((h, w) + ((Cn,) if Cn > 1 else ())))
            off += n
        return out

    def depth_map_pyramid(self, depth, ratio, min_width):
        d = _f64(depth)
        H, W = d.shape
        n = len(pyramid_levels(W, H, ratio, min_width))
        r = ratio if 0.4 <= ratio <= 0.98 else 0.75
        sizes = [(int(W * r ** m), int(H * r ** m)) for m in range(n)]
        flat = np.empty(sum(w * h for w, h in sizes), np.float64)
        self._check(lib().rgbdflow_depth_map_pyramid(self.h, _p(d), W, H, C.c_double(ratio), min_width, _p(flat)))
        out, off = [], 0
        for (w, h) in sizes:
            out.append(flat[off:off + w * h].reshape(h, w))
            off += w * h
        return out

    def fill_depth_holes(self, depth):
        """The device's parallel hole fill of a (H, W) depth map (metres, <= 0
        invalid): (filled f64 map, pixels left invalid)."""
        d = _f64(depth)
        H, W = d.shape
        out = np.empty((H, W))
        left = C.c_int()
        self._check(lib().rgbdflow_fill_depth_holes(self.h, _p(d), W, H, _p(out), C.byref(left)))
        return out, left.value

    def fill_depth_holes_noalloc(self, depth_ptr, W, H, out_ptr, unfilled_ptr=None, stream=None):
        """rgbdflow_fill_depth_holes_noalloc on device pointers (ints): enqueues
        the fill on `stream` (a hipStream_t as an int, None: the context's
        stream) and returns without waiting."""
        self._check(lib().rgbdflow_fill_depth_holes_noalloc(
            self.h, C.c_void_p(stream) if stream else None, C.c_void_p(depth_ptr), int(W), int(H),
            C.c_void_p(out_ptr), C.c_void_p(unfilled_ptr) if unfilled_ptr else None))

    def fill_small_holes_colour(self, depth, bgr, max_region=60):
        """The device's colour-guided fill of small holes: (filled f64 map,
        pixels of small regions left invalid).  bgr: (H, W, 3) in [0, 1]."""
        d, b = _f64(depth), _f64(bgr)
        H, W = d.shape
        out = np.empty((H, W))
        left = C.c_int()
        self._check(lib().rgbdflow_fill_small_holes_colour(self.h, _p(d), _p(b), W, H, max_region, _p(out),
                                                          C.byref(left)))
        return out, left.value

    def heuristic_pb(self, bgr, depth):
        """computeHeuristicPb of one image on the device: (H, W, 3) BGR in [0,1],
        (H, W) metres -> (H, W, 2) f32."""
        b, d = _f64(bgr), _f64(depth)
        H, W = d.shape
        out = np.empty((H, W, 2), np.float32)
        self._check(lib().rgbdflow_heuristic_pb(self.h, _p(b), _p(d), W, H, _p(out)))
        return out

    def im2feature(self, bgr):
        bgr = _f64(bgr)
        H, W = bgr.shape[:2]
        out = np.empty((H, W, 5), np.float64)
        self._check(lib().rgbdflow_im2feature(self.h, _p(bgr), W, H, _p(out)))
        return out

    def warp_by_3d_flow(self, I1, D1, I2, D2, u, v, w):
        I1, D1, I2, D2, u, v, w = map(_f64, (I1, D1, I2, D2, u, v, w))
        H, W = D1.shape
        wI = np.empty_like(I1)
        wD = np.empty_like(D1)
        self._check(lib().rgbdflow_warp_by_3d_flow(self.h, _p(I1), _p(D1), _p(I2), _p(D2), _p(u), _p(v), _p(w), W,
                                                   H, I1.shape[2], _p(wI), _p(wD)))
        return wI, wD

    def upsample_flow(self, u, v, w, Wd, Hd, uv_scale):
        u, v, w = _f64(u), _f64(v), _f64(w)
        Hs, Ws = u.shape
        uo = np.empty((Hd, Wd)); vo = np.empty((Hd, Wd)); wo = np.empty((Hd, Wd))
        self._check(lib().rgbdflow_upsample_flow(self.h, _p(u), _p(v), _p(w), Ws, Hs, Wd, Hd, C.c_double(uv_scale),
                                                 _p(uo), _p(vo), _p(wo)))
        return uo, vo, wo

    def depth_stdevs(self, depth_f32, max_dist=4):
        d = _f32(depth_f32)
        H, W = d.shape
        out = np.empty_like(d)
        self._check(lib().rgbdflow_depth_stdevs_pointwise(self.h, _p(d), W, H, max_dist, _p(out)))
        return out

    @staticmethod
    def _level_args(li):
        return (C.c_float(li.focal_level), _p(li.I1), _p(li.D1), _p(li.I2), _p(li.D2), _p(li.warpI2),
                _p(li.warpD2))

    def level_system(self, li, params, u, v, w):
        u, v, w = _f64(u), _f64(v), _f64(w)
        out = np.empty((13, li.H, li.W), np.float64)
        self._check(lib().rgbdflow_level_system(
            self.h, li.W, li.H, *self._level_args(li), _p(u), _p(v), _p(w), C.c_double(params.color_data_weight),
            _p(li.delta), C.c_double(params.alpha), _p(li.s), C.c_double(params.gamma), _p(li.m), _p(li.r),
            _p(out)))
        return out

    def relax(self, coef13, n_iter, x, schedule: Params | None = None):
        c = _f64(coef13)
        _, H, W = c.shape
        xx = _f64(x).copy()
        self._check(lib().rgbdflow_relax(self.h, _p(c), W, H, n_iter, C.byref(schedule) if schedule else None,
                                         _p(xx)))
        return xx

    def relax_residual(self, coef13, max_iter, tol, x, schedule: Params | None = None):
        """Relax from x until the FP64 relative residual (checked every 25
        iterations) is <= tol or max_iter iterations ran; returns
        (x, iterations, residual, energy)."""
        c = _f64(coef13)
        _, H, W = c.shape
        xx = _f64(x).copy()
        it, res, en = C.c_int(), C.c_double(), C.c_double()
        self._check(lib().rgbdflow_relax_residual(self.h, _p(c), W, H, int(max_iter), C.c_double(tol),
                                                  C.byref(schedule) if schedule else None, _p(xx), C.byref(it),
                                                  C.byref(res), C.byref(en)))
        return xx, it.value, res.value, en.value

    def solve_level(self, li, params, allow_early, u, v, w):
        u, v, w = _f64(u).copy(), _f64(v).copy(), _f64(w).copy()
        st = Stats()
        self._check(lib().rgbdflow_solve_level(
            self.h, li.W, li.H, *self._level_args(li), _p(u), _p(v), _p(w), C.c_double(params.color_data_weight),
            _p(li.delta), C.c_double(params.alpha), _p(li.s), C.c_double(params.gamma), _p(li.m), _p(li.r),
            params.n_outer, params.n_inner, params.n_sor, int(allow_early), C.byref(params), C.byref(st)))
        return u, v, w, st

    # ------------------------------------------------------------ pair path
    def pair_load(self, params, im1, dm1, im2, dm2, sigma1=None, hpb=None):
        im1, dm1, im2, dm2 = _f64(im1), _f64(dm1), _f64(im2), _f64(dm2)
        H, W = dm1.shape
        s1 = _f32(sigma1) if sigma1 is not None else None
        h = _f32(hpb) if hpb is not None else None
        self._keep = (im1, dm1, im2, dm2, s1, h)
        self._check(lib().rgbdflow_pair_load(self.h, C.byref(params), W, H, _p(im1), _p(dm1), _p(im2), _p(dm2),
                                             _p(s1), _p(h)))
        self._pair_shape = (H, W)
        self._batch_shapes = [(H, W)]

    def pair_run(self, params, stats=True):
        if stats:
            st = Stats()
            self._check(lib().rgbdflow_pair_run(self.h, C.byref(params), C.byref(st)))
            return st
        self._check(lib().rgbdflow_pair_run(self.h, C.byref(params), None))
        return None

    def pair_read(self):
        H, W = self._pair_shape
        u = np.empty((H, W)); v = np.empty((H, W)); w = np.empty((H, W))
        self._check(lib().rgbdflow_pair_read(self.h, _p(u), _p(v), _p(w)))
        return u, v, w

    # ---------------------------------------------------------------- batches
    def batch_load(self, params, pairs):
        """pairs: list of dicts with im1, dm1, im2, dm2 and optionally sigma1,
        hpb, cam (a Camera), min_width, est (an (H, W, 3) flow estimate: the
        pair is then solved as sceneFlowWithInit) and hpb_dm1 (the depth the
        boundary maps made on the device start from)."""
        keep, arr = [], (PairInput * len(pairs))()
        for i, pr in enumerate(pairs):
            im1, dm1, im2, dm2 = _f64(pr["im1"]), _f64(pr["dm1"]), _f64(pr["im2"]), _f64(pr["dm2"])
            s1 = _f32(pr["sigma1"]) if pr.get("sigma1") is not None else None
            h = _f32(pr["hpb"]) if pr.get("hpb") is not None else None
            est = _f64(pr["est"]) if pr.get("est") is not None else None
            hd = _f64(pr["hpb_dm1"]) if pr.get("hpb_dm1") is not None else None
            keep.append((im1, dm1, im2, dm2, s1, h, est, hd))
            H, W = dm1.shape
            a = arr[i]
            a.W, a.H = W, H
            a.im1, a.dm1, a.im2, a.dm2 = _p(im1), _p(dm1), _p(im2), _p(dm2)
            a.sigma1, a.hpb, a.est, a.hpb_dm1 = _p(s1), _p(h), _p(est), _p(hd)
            if pr.get("cam") is not None:
                a.cam = pr["cam"]
            a.min_width = int(pr.get("min_width", 0))
        self._check(lib().rgbdflow_batch_load(self.h, C.byref(params), len(pairs), arr))
        self._batch_shapes = [k[1].shape for k in keep]
        self._pair_shape = self._batch_shapes[0]

    def batch_run(self, params, stats=True):
        n = len(self._batch_shapes)
        if stats:
            st = (Stats * n)()
            self._check(lib().rgbdflow_batch_run(self.h, C.byref(params), st))
            return list(st)
        self._check(lib().rgbdflow_batch_run(self.h, C.byref(params), None))
        return None

    def batch_read(self, i):
        H, W = self._batch_shapes[i]
        u = np.empty((H, W)); v = np.empty((H, W)); w = np.empty((H, W))
        self._check(lib().rgbdflow_batch_read(self.h, int(i), _p(u), _p(v), _p(w)))
        return u, v, w

    def multiscale(self, params, im1, dm1, im2, dm2, sigma1=None, hpb=None):
        self.pair_load(params, im1, dm1, im2, dm2, sigma1, hpb)
        st = self.pair_run(params, stats=True)
        u, v, w = self.pair_read()
        return u, v, w, st

    def with_init(self, params, im1, dm1, im2, dm2, est, sigma1=None, hpb0=None):
        im1, dm1, im2, dm2, est = _f64(im1), _f64(dm1), _f64(im2), _f64(dm2), _f64(est)
        H, W = dm1.shape
        s1 = _f32(sigma1) if sigma1 is not None else None
        h = _f32(hpb0) if hpb0 is not None else None
        u = np.empty((H, W)); v = np.empty((H, W)); w = np.empty((H, W))
        st = Stats()
        self._check(lib().rgbdflow_with_init(self.h, W, H, _p(im1), _p(dm1), _p(im2), _p(dm2), _p(s1), _p(h),
                                             _p(est), C.byref(params), _p(u), _p(v), _p(w), C.byref(st)))
        return u, v, w, st


class Server:
    """A serving runtime (rgbdflow_server_*): streams of frames, continuous batching."""

    def __init__(self, params, device=0, max_batch=16, lanes=2, hbm_budget=0, warm_start=False, warm_n_outer=0,
                 fill_holes=False):
        """warm_start: every pair after a stream's first is solved as
        sceneFlowWithInit from the stream's previous result (warm_n_outer > 0:
        with that many outer iterations).  fill_holes: each frame's depth is
        hole-filled on the device (depth stdevs from the depth as submitted)."""
        cfg = ServerConfig(device, max_batch, lanes, int(warm_start), hbm_budget, warm_n_outer, int(fill_holes))
        h = C.c_void_p()
        rc = lib().rgbdflow_server_create(C.byref(h), C.byref(cfg), C.byref(params))
        if rc != OK:
            raise RgbdFlowError(rc, "rgbdflow_server_create failed")
        self.h = h
        self._sizes = {}

    def _check(self, rc):
        if rc < 0:
            raise RgbdFlowError(rc, lib().rgbdflow_server_last_error(self.h).decode())
        return rc

    def open(self, W, H, cam=None, min_width=0):
        sid = C.c_int()
        self._check(lib().rgbdflow_stream_open(self.h, W, H, C.byref(cam) if cam is not None else None, min_width,
                                               C.byref(sid)))
        self._sizes[sid.value] = (H, W)
        return sid.value

    def submit(self, stream, bgr_u8, depth_f32, hpb=None):
        b = np.ascontiguousarray(bgr_u8, dtype=np.uint8)
        d = _f32(depth_f32)
        hh = _f32(hpb) if hpb is not None else None
        t = C.c_ulonglong()
        self._check(lib().rgbdflow_stream_submit(self.h, stream, _p(b), _p(d), _p(hh), C.byref(t)))
        return (t.value, stream)

    def poll(self, ticket, wait_ms=-1):
        """(info dict, (u, v, w) or None) when done; None if not done in wait_ms."""
        t, stream = ticket
        H, W = self._sizes[stream]
        u = np.empty((H, W)); v = np.empty((H, W)); w = np.empty((H, W))
        info = RequestInfo()
        rc = self._check(lib().rgbdflow_stream_poll(self.h, C.c_ulonglong(t), int(wait_ms), _p(u), _p(v), _p(w),
                                                    C.byref(info)))
        if rc == 0:
            return None
        return info.as_dict(), ((u, v, w) if info.has_flow else None)

    def set_relax_schedule(self, W, H, choice: dict):
        ch = TileChoice()
        for f, _ in TileChoice._fields_:
            if f in choice:
                setattr(ch, f, choice[f])
        self._check(lib().rgbdflow_server_set_relax_schedule(self.h, W, H, C.byref(ch)))

    def close_stream(self, stream):
        self._check(lib().rgbdflow_stream_close(self.h, stream))

    def close(self):
        if self.h:
            lib().rgbdflow_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
