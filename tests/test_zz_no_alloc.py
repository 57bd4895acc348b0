"""rgbdflow_fill_depth_holes_noalloc (device memory in and out, scratch in a
fixed per-device array, nothing allocated) against rgbdflow_fill_depth_holes
on the same frames: bit for bit, on two calls in a row with the second
smaller, on the caller's stream and on the context's, out of place and in
place; a frame above the limit is refused before anything is enqueued."""
import ctypes as C

import numpy as np
import pytest

from helpers import load_package

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip(R):
    """The HIP runtime librgbdflow.so itself uses (device memory and streams
    of the test come from it, not from another runtime in the process)."""
    R.lib()
    paths = [ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64.so" in ln and "/torch/" not in ln]
    assert paths, "librgbdflow.so's HIP runtime is not mapped"
    hip = C.CDLL(paths[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


class _Dev:
    """A device buffer holding a copy of a numpy array."""

    def __init__(self, hip, a):
        self.hip, self.shape, self.dtype = hip, a.shape, a.dtype
        self.nbytes = a.nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), self.nbytes) == 0
        self.put(a)

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, self.nbytes, H2D) == 0

    def get(self):
        a = np.empty(self.shape, self.dtype)
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, self.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _frame(W, H, seed, holes):
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    d = 1.5 + 0.002 * xs - 0.001 * ys + rng.normal(0, 1e-3, (H, W))
    d[rng.random((H, W)) < holes] = 0.0
    d[H // 4:H // 4 + H // 8, W // 5:W // 5 + W // 6] = 0.0
    d[:, :3] = -1.0
    return d


def test_noalloc_fill_equals_the_allocating_entry_point():
    R = load_package()
    hip = _hip(R)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    bufs = []
    try:
        with R.Context(0) as ctx:
            # first call: 640x480, 5% holes, a block and a border strip, on the caller's stream, out of place
            d1 = _frame(640, 480, 11, 0.05)
            want1, left1 = ctx.fill_depth_holes(d1)
            print(f"no-alloc fill 640x480: {(d1 <= 0).sum()} invalid, caller's stream, out of place", flush=True)
            src = _Dev(hip, d1)
            out = _Dev(hip, np.full_like(d1, 7.0))
            left = _Dev(hip, np.array([-5], np.int32))
            bufs += [src, out, left]
            ctx.fill_depth_holes_noalloc(src.p.value, 640, 480, out.p.value, left.p.value, stream=stream.value)
            assert hip.hipStreamSynchronize(stream) == 0
            assert np.array_equal(out.get(), want1)
            assert int(left.get()[0]) == left1
            assert np.array_equal(src.get(), d1)  # the input is left as it was

            # second call, smaller: 97x61 with 90% holes (many rounds), in place, on the context's stream
            d2 = _frame(97, 61, 12, 0.9)
            want2, left2 = ctx.fill_depth_holes(d2)
            print(f"no-alloc fill 97x61: {(d2 <= 0).sum()} invalid, context stream, in place", flush=True)
            buf = _Dev(hip, d2)
            bufs.append(buf)
            left.put(np.array([-5], np.int32))
            ctx.fill_depth_holes_noalloc(buf.p.value, 97, 61, buf.p.value, left.p.value)
            ctx.synchronize()
            assert np.array_equal(buf.get(), want2)
            assert int(left.get()[0]) == left2

            # the count is optional
            buf.put(d2)
            ctx.fill_depth_holes_noalloc(buf.p.value, 97, 61, buf.p.value)
            ctx.synchronize()
            assert np.array_equal(buf.get(), want2)

            # above the limit of the fixed scratch: refused on the host, nothing enqueued
            big = _Dev(hip, np.zeros(1088 * 1921))
            bufs.append(big)
            with pytest.raises(R.RgbdFlowError) as e:
                ctx.fill_depth_holes_noalloc(big.p.value, 1921, 1088, big.p.value)
            assert e.value.code == R.ERR_ARG and "limit" in str(e.value)
            ctx.synchronize()
    finally:
        for b in bufs:
            b.free()
        hip.hipStreamDestroy(stream)
