"""ampcor_hip_adjust_noalloc through the C ABI against the plan path.

The allocation-free entry point runs the same kernels with the same (untuned)
tilings as a plan, its buffers carved from the library's fixed scratch array,
so its products must equal the plan's bit for bit -- on two calls in a row,
the second on fewer pairs (the array keeps the first call's contents).  A
batch larger than the array is refused with a clear error.  The test allocates the rasters' and
products' device memory itself, with the HIP runtime.
"""

import ctypes

import numpy as np
import pytest

from ampcor_amd import build, synth
from tests.gpu_helpers import run_hip

pytestmark = pytest.mark.gpu

PRECISION = {"fp32": 0, "bf16": 1}


@pytest.fixture(scope="module")
def abi(lib):
    L = ctypes.CDLL(build.library_path())
    sz, vp, i64p = ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
    L.ampcor_hip_adjust_noalloc.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, sz, sz, vp, sz, sz,
                                            sz, i64p, i64p, sz, sz, sz, sz, sz, ctypes.c_int,
                                            vp, vp, vp, vp]
    L.ampcor_hip_adjust_noalloc.restype = ctypes.c_int
    L.ampcor_hip_last_error.restype = ctypes.c_char_p
    return L


class Hip:
    """The few HIP runtime calls the test needs for its own device buffers."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        self.L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        self.L.hipFree.argtypes = [vp]
        self.L.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.malloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return p

    def download(self, p, shape):
        out = np.empty(shape, np.float32)
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0
        return out


def noalloc(abi, hip, s, n, precision):
    ref, tgt = hip.upload(s.reference), hip.upload(s.target)
    shapes = {"offsets": (n, 2), "coarse_peak": (n,), "refined_peak": (n,), "snr": (n,)}
    out = {k: hip.upload(np.full(v, np.nan, np.float32)) for k, v in shapes.items()}
    rb = np.ascontiguousarray(s.ref_begin[:n], dtype=np.int64)
    tb = np.ascontiguousarray(s.tgt_begin[:n], dtype=np.int64)
    i64p = ctypes.POINTER(ctypes.c_int64)
    try:
        rc = abi.ampcor_hip_adjust_noalloc(
            0, None, 0, ref, *s.reference.shape, tgt, *s.target.shape, n,
            rb.ctypes.data_as(i64p), tb.ctypes.data_as(i64p), s.chip, s.window, s.refine_margin,
            s.refine_factor, s.zoom_factor, PRECISION[precision], out["offsets"],
            out["coarse_peak"], out["refined_peak"], out["snr"])
        assert rc == 0, abi.ampcor_hip_last_error()
        return {k: hip.download(out[k], shapes[k]) for k in shapes}
    finally:
        for p in [ref, tgt, *out.values()]:
            hip.L.hipFree(p)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_noalloc_equals_plan_adjust(lib, abi, precision):
    hip = Hip()
    s = synth.shifted_scene(seed=12, size=512, grid=(8, 8), chip=32, margin=16, zoom_factor=4,
                            shift=(2.31, -1.62))
    for n in (s.pairs, 23):   # the second call on fewer pairs
        print(f"noalloc {precision}: {n} pairs", flush=True)
        got = noalloc(abi, hip, s, n, precision)
        want = run_hip(lib, s.reference, s.target, s.ref_begin[:n], s.tgt_begin[:n], s.chip,
                       s.window, s.refine_factor, s.refine_margin, s.zoom_factor,
                       precision=precision)
        for k in ("offsets", "coarse_peak", "refined_peak", "snr"):
            assert np.array_equal(got[k].view(np.uint32), np.asarray(want[k]).view(np.uint32)), k


def test_noalloc_refuses_batches_beyond_the_scratch(lib, abi):
    s = synth.shifted_scene(seed=12, size=512, grid=(8, 8), chip=32, margin=16, zoom_factor=4)
    s.ref_begin = np.repeat(s.ref_begin, 200, axis=0)
    s.tgt_begin = np.repeat(s.tgt_begin, 200, axis=0)
    n = len(s.ref_begin)   # 12800 pairs: far beyond 64 MiB at this geometry
    i64p = ctypes.POINTER(ctypes.c_int64)
    rb = np.ascontiguousarray(s.ref_begin, dtype=np.int64)
    tb = np.ascontiguousarray(s.tgt_begin, dtype=np.int64)
    one = ctypes.c_void_p(1)   # never dereferenced: the size check comes first
    rc = abi.ampcor_hip_adjust_noalloc(0, None, 0, one, *s.reference.shape, one,
                                       *s.target.shape, n, rb.ctypes.data_as(i64p),
                                       tb.ctypes.data_as(i64p), s.chip, s.window, s.refine_margin,
                                       s.refine_factor, s.zoom_factor, 0, one, one, one, one)
    msg = abi.ampcor_hip_last_error().decode()
    print(msg, flush=True)
    assert rc == 1 and "split the batch" in msg and "at most" in msg
