"""Helpers shared by the GPU tests: run a scene through the C ABI module."""

import numpy as np


def run_hip(lib, reference, target, ref_begin, tgt_begin, chip, window, rf, rm, zf,
            per_pair=False, precision="fp32", tune=False):
    n = len(ref_begin)
    worker = lib.sequential(n, (chip, chip), (window, window), rm, rf, zf, precision=precision)
    dtype = "complex64" if reference.dtype == np.complex64 else "float32"
    ref = lib.raster(np.ascontiguousarray(reference), *reference.shape, dtype=dtype)
    tgt = lib.raster(np.ascontiguousarray(target), *target.shape, dtype=dtype)
    rb = np.ascontiguousarray(ref_begin, dtype=np.int64)
    tb = np.ascontiguousarray(tgt_begin, dtype=np.int64)
    if per_pair:
        for k in range(n):
            r0, t0 = tuple(int(v) for v in rb[k]), tuple(int(v) for v in tb[k])
            lib.addReference(worker, ref, k, r0, (r0[0] + chip, r0[1] + chip))
            lib.addTarget(worker, tgt, k, t0, (t0[0] + window, t0[1] + window))
    else:
        lib.addPairs(worker, ref, tgt, 0, rb.tobytes(), tb.tobytes())
    if tune:   # the autotuner times every candidate tiling on this data, then keeps the fastest
        lib.adjust(worker, offsets=False)
        lib.autotune(worker, 1)
    offsets = np.array(lib.adjust(worker), dtype=np.float32).reshape(n, 2)
    P, R, Z = lib.dims(worker)
    prods = {k: np.frombuffer(v, np.float32) for k, v in lib.products(worker).items()}
    locs = {k: np.frombuffer(v, np.int32).reshape(n, 2) for k, v in lib.stage_locations(worker).items()}
    out = {
        "tuple_offsets": offsets,
        "offsets": prods["offsets"].reshape(n, 2),
        "coarse_peak": prods["coarse_peak"],
        "refined_peak": prods["refined_peak"],
        "snr": prods["snr"],
        "coarse_loc": locs["coarse"],
        "window_loc": locs["window"],
        "zoom_loc": locs["zoomed"],
        "gamma": np.frombuffer(lib.stage_gamma(worker), np.float32).reshape(n, P, P),
        "gamma_refined": np.frombuffer(lib.stage_gamma_refined(worker), np.float32).reshape(n, R, R),
        "guard": lib.fp8_guard(worker),
    }
    return out
