"""The C ABI library and the CPython module load and validate without a GPU."""

import ctypes
import os
import re

import numpy as np
import pytest

from ampcor_amd import build, ext

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ampcor_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(ampcor_hip_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 18
    lib = ctypes.CDLL(build.library_path())
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_module_mirrors_the_cuda_extension_entry_points():
    lib = ext.require()
    for name in ("sequential", "addReference", "addTarget", "adjust", "refine", "slc_map",
                 "slc_pixelSize", "slc_getitem"):
        assert callable(getattr(lib, name))
    assert lib.slc_pixelSize() == 8
    # invalid capsules raise TypeError, as in ext/ampcor/sequential.cc:277-283 and slc.cc:124-129
    with pytest.raises(TypeError):
        lib.refine(object())
    with pytest.raises(TypeError):
        lib.slc_getitem(object(), 0)
    assert issubclass(lib.Error, Exception)


def test_plan_arguments_are_validated_before_any_device_work():
    lib = ext.require()
    with pytest.raises(ValueError, match="square"):
        lib.sequential(4, (32, 16), (64, 64), 8, 2, 4)
    with pytest.raises(ValueError, match="refine margin"):
        lib.sequential(4, (32, 32), (64, 64), 20, 2, 4)
    with pytest.raises(ValueError, match="even"):
        lib.sequential(4, (32, 32), (65, 65), 8, 2, 4)
    with pytest.raises(ValueError, match="precision"):
        lib.sequential(4, (32, 32), (64, 64), 8, 2, 4, precision="fp7")
    with pytest.raises(ValueError, match="at least one pair"):
        lib.sequential(0, (32, 32), (64, 64), 8, 2, 4)


def test_c_abi_status_codes_without_a_device():
    L = ctypes.CDLL(build.library_path())
    L.ampcor_hip_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    rc = L.ampcor_hip_plan_create(0, ctypes.c_size_t(4), ctypes.c_size_t(32), ctypes.c_size_t(32),
                                  ctypes.c_size_t(64), ctypes.c_size_t(64), ctypes.c_size_t(8),
                                  ctypes.c_size_t(2), ctypes.c_size_t(4), 7, ctypes.byref(out))
    assert rc == 1 and b"FP32" in L.ampcor_hip_last_error()
    assert L.ampcor_hip_adjust(None, None) == 1
    assert L.ampcor_hip_plan_destroy(None) == 0
    assert L.ampcor_hip_raster_destroy(None) == 0
    assert L.ampcor_hip_raster_read(None, ctypes.c_size_t(0), ctypes.c_size_t(1), None) == 1
    n = ctypes.c_int(-1)
    assert L.ampcor_hip_device_count(ctypes.byref(n)) == 0 and n.value >= 0


def test_geometry_the_fp8_form_cannot_run_is_refused_without_a_device():
    # the predicate a server's submit shares with plan creation: FP8 needs
    # (window - chip + 1) <= 144, which chip 32 / window 320 breaks
    L = ctypes.CDLL(build.library_path())
    L.ampcor_hip_last_error.restype = ctypes.c_char_p
    sz, out = ctypes.c_size_t, ctypes.c_void_p()
    rc = L.ampcor_hip_plan_create(0, sz(4), sz(32), sz(32), sz(320), sz(320), sz(8), sz(2), sz(8),
                                  5, ctypes.byref(out))
    assert rc == 1 and b"FP8" in L.ampcor_hip_last_error()



def test_invalid_capsules_raise_type_error():
    lib = ext.require()
    with pytest.raises(TypeError):
        lib.adjust(object())
    import ctypes as c
    bogus = c.pythonapi.PyCapsule_New
    bogus.restype = c.py_object
    bogus.argtypes = [c.c_void_p, c.c_char_p, c.c_void_p]
    cap = bogus(1, b"something::else", None)
    with pytest.raises(TypeError, match="invalid HIP Sequential worker capsule"):
        lib.adjust(cap)


def test_raster_buffer_size_is_checked():
    lib = ext.require()
    with pytest.raises(ValueError, match="bytes"):
        lib.raster(np.zeros((4, 4), np.complex64), 4, 5)


def test_server_needs_a_device_and_valid_arguments():
    lib = ext.require()
    with pytest.raises(ValueError):
        lib.server(batch=0)
    if lib.device_count() == 0:
        with pytest.raises(lib.Error, match="no HIP device"):
            lib.server()


def test_serving_jobs_stay_inside_their_rasters():
    from ampcor_amd import serving
    pool = [(np.zeros((300, 260), np.complex64), None), (np.zeros((512, 512), np.complex64), None)]
    for j in serving.make_jobs(pool, count=40, rate=50.0, seed=1, min_pairs=8, max_pairs=64):
        lines, samples = pool[j.raster][0].shape
        assert np.all(j.tgt_begin >= 0)
        assert np.all(j.tgt_begin[:, 0] + j.window <= lines)
        assert np.all(j.tgt_begin[:, 1] + j.window <= samples)
        assert np.all(j.ref_begin - j.tgt_begin == j.margin)
        assert 8 <= j.pairs <= 64
