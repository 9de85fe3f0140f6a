"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

FORMS = {"fp64": 0, "fp32": 1}
DTYPES = {np.dtype(np.complex64): 0, np.dtype(np.float32): 1}

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i = ctypes.POINTER(ctypes.c_int32)


class Geometry(ctypes.Structure):
    _fields_ = [("chip", ctypes.c_int), ("window", ctypes.c_int), ("refine_factor", ctypes.c_int),
                ("refine_margin", ctypes.c_int), ("zoom_factor", ctypes.c_int)]


class Products(ctypes.Structure):
    _fields_ = [("offsets", _f), ("coarse_peak", _f), ("refined_peak", _f), ("snr", _f),
                ("coarse_loc", _i), ("window_loc", _i), ("zoom_loc", _i), ("gamma", _f),
                ("gamma_refined", _f)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        c_int, c_i64, vp = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p
        L.oracle_sat.argtypes = [_d, c_int, c_int, _d]
        L.oracle_sat_sum.argtypes = [_d, c_int, c_int, c_int, c_int, c_int]
        L.oracle_sat_sum.restype = ctypes.c_double
        L.oracle_ncc64.argtypes = [_d, c_int, _d, c_int, _d]
        L.oracle_cross_sum.argtypes = [_d, c_int, _d, c_int, _d]
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_argmax_f32.argtypes = [_f, c_int, c_int, ip, ip, _f]
        L.oracle_argmax_f64.argtypes = [_d, c_int, c_int, ip, ip, _d]
        L.oracle_nudge.argtypes = [c_int, c_int, c_int, c_int, c_int, ip, ip]
        L.oracle_upsample.argtypes = [_d, c_int, c_int, _d]
        L.oracle_upsample_kernel.argtypes = [c_int, c_int, c_int, c_int]
        L.oracle_upsample_kernel.restype = ctypes.c_double
        L.oracle_zoom_full.argtypes = [_d, c_int, c_int, _d]
        L.oracle_dirichlet.argtypes = [c_int, c_int, c_int]
        L.oracle_dirichlet.restype = ctypes.c_double
        L.oracle_parabola.argtypes = [ctypes.c_double] * 3
        L.oracle_parabola.restype = ctypes.c_double
        L.oracle_zoom_search.argtypes = [_d, c_int, c_int, ip, ip, _d, _d, _d]
        L.oracle_zoom_search.restype = c_int
        L.oracle_snr.argtypes = [_d, c_int, c_int, c_int, c_int, c_int]
        L.oracle_snr.restype = ctypes.c_double
        L.oracle_adjust.argtypes = [c_int, c_int, vp, c_i64, c_i64, vp, c_i64, c_i64, c_i64, _i, _i,
                                    Geometry, c_int, Products]
        L.oracle_adjust.restype = c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def sat(t):
    t = np.ascontiguousarray(t, dtype=np.float64)
    r, c = t.shape
    S = np.empty((r + 1, c + 1))
    lib().oracle_sat(_p(t, _d), r, c, _p(S, _d))
    return S


def sat_sum(S, r0, c0, r1, c1):
    S = np.ascontiguousarray(S, dtype=np.float64)
    return lib().oracle_sat_sum(_p(S, _d), S.shape[1] - 1, r0, c0, r1, c1)


def ncc64(ref, tgt):
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    tgt = np.ascontiguousarray(tgt, dtype=np.float64)
    C, W = ref.shape[0], tgt.shape[0]
    g = np.empty((W - C + 1, W - C + 1))
    lib().oracle_ncc64(_p(ref, _d), C, _p(tgt, _d), W, _p(g, _d))
    return g


def cross_sum(ref, tgt):
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    tgt = np.ascontiguousarray(tgt, dtype=np.float64)
    C, W = ref.shape[0], tgt.shape[0]
    g = np.empty((W - C + 1, W - C + 1))
    lib().oracle_cross_sum(_p(ref, _d), C, _p(tgt, _d), W, _p(g, _d))
    return g


def argmax(g):
    r, c = ctypes.c_int(), ctypes.c_int()
    if g.dtype == np.float32:
        g = np.ascontiguousarray(g)
        pk = ctypes.c_float()
        lib().oracle_argmax_f32(_p(g, _f), g.shape[0], g.shape[1], ctypes.byref(r), ctypes.byref(c),
                                ctypes.byref(pk))
    else:
        g = np.ascontiguousarray(g, dtype=np.float64)
        pk = ctypes.c_double()
        lib().oracle_argmax_f64(_p(g, _d), g.shape[0], g.shape[1], ctypes.byref(r), ctypes.byref(c),
                                ctypes.byref(pk))
    return r.value, c.value, pk.value


def nudge(row, col, refine_margin, window, expanded):
    t, l = ctypes.c_int(), ctypes.c_int()
    lib().oracle_nudge(row, col, refine_margin, window, expanded, ctypes.byref(t), ctypes.byref(l))
    return t.value, l.value


def upsample(x, rf):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    N = x.shape[0]
    out = np.empty((rf * N, rf * N), dtype=np.complex128)
    lib().oracle_upsample(_p(x.view(np.float64), _d), N, rf, _p(out.view(np.float64), _d))
    return out


def upsample_kernel(N, rf, p, n):
    return lib().oracle_upsample_kernel(N, rf, p, n)


def dirichlet(R, zf, d):
    return lib().oracle_dirichlet(R, zf, d)


def zoom_full(g, zf):
    g = np.ascontiguousarray(g, dtype=np.float64)
    R = g.shape[0]
    out = np.empty((zf * R, zf * R))
    lib().oracle_zoom_full(_p(g, _d), R, zf, _p(out, _d))
    return out


def parabola(fm, f0, fp):
    return lib().oracle_parabola(fm, f0, fp)


def zoom_search(g, zf):
    """The pipeline's zoom search on one refined surface: ((p, q), peak, (dp, dq))."""
    g = np.ascontiguousarray(g, dtype=np.float64)
    p, q = ctypes.c_int(), ctypes.c_int()
    pk, dp, dq = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = lib().oracle_zoom_search(_p(g, _d), g.shape[0], zf, ctypes.byref(p), ctypes.byref(q),
                                  ctypes.byref(pk), ctypes.byref(dp), ctypes.byref(dq))
    if rc:
        raise ValueError("zoom search rejected its arguments")
    return (p.value, q.value), pk.value, (dp.value, dq.value)


def snr(g, prow, pcol, exclude):
    g = np.ascontiguousarray(g, dtype=np.float64)
    return lib().oracle_snr(_p(g, _d), g.shape[0], g.shape[1], prow, pcol, exclude)


def adjust(reference, target, ref_begin, tgt_begin, chip, window, refine_factor, refine_margin,
           zoom_factor, form="fp64", threads=0, surfaces=False):
    """Run the whole pipeline; returns a dict of numpy arrays (pair-major)."""
    reference = np.ascontiguousarray(reference)
    target = np.ascontiguousarray(target)
    if reference.dtype != target.dtype or reference.dtype not in DTYPES:
        raise TypeError("rasters must both be complex64 or float32")
    rb = np.ascontiguousarray(ref_begin, dtype=np.int32).reshape(-1, 2)
    tb = np.ascontiguousarray(tgt_begin, dtype=np.int32).reshape(-1, 2)
    n = rb.shape[0]
    P = window - chip + 1
    R = 2 * refine_factor * refine_margin + 1
    out = {
        "offsets": np.empty((n, 2), np.float32), "coarse_peak": np.empty(n, np.float32),
        "refined_peak": np.empty(n, np.float32), "snr": np.empty(n, np.float32),
        "coarse_loc": np.empty((n, 2), np.int32), "window_loc": np.empty((n, 2), np.int32),
        "zoom_loc": np.empty((n, 2), np.int32),
    }
    prods = Products(_p(out["offsets"], _f), _p(out["coarse_peak"], _f), _p(out["refined_peak"], _f),
                     _p(out["snr"], _f), _p(out["coarse_loc"], _i), _p(out["window_loc"], _i),
                     _p(out["zoom_loc"], _i), None, None)
    if surfaces:
        out["gamma"] = np.empty((n, P, P), np.float32)
        out["gamma_refined"] = np.empty((n, R, R), np.float32)
        prods.gamma = _p(out["gamma"], _f)
        prods.gamma_refined = _p(out["gamma_refined"], _f)
    geom = Geometry(chip, window, refine_factor, refine_margin, zoom_factor)
    rc = lib().oracle_adjust(FORMS[form], DTYPES[reference.dtype], reference.ctypes.data,
                             reference.shape[0], reference.shape[1], target.ctypes.data,
                             target.shape[0], target.shape[1], n, _p(rb, _i), _p(tb, _i), geom,
                             threads, prods)
    if rc:
        raise ValueError(f"oracle_adjust rejected its arguments (status {rc})")
    return out
