"""The raster argument of addReference / addTarget / addPairs, as the reference
hands it over (ext/ampcor_cuda/sequential.cc:85-143 takes the raster that
SLC.open made, ext/ampcor/slc.cc:50-93 maps it from a raw complex64 file).

The same scene reaches the worker four ways -- a raw complex64 file mapped by
slc_map, an explicit raster() upload, the NumPy arrays themselves, and the
pyre-free SLC wrapper -- and the offsets, peaks and SNR are bitwise equal.
The CPU module's SLC capsule and malformed arrays are refused with errors
that name the replacement.
"""

import ctypes
import sys

import numpy as np
import pytest

from ampcor_amd import dom, synth

pytestmark = pytest.mark.gpu

CHIP, WINDOW, RF, RM, ZF = 32, 64, 2, 8, 8


@pytest.fixture(scope="module")
def scene():
    return synth.shifted_scene(seed=11, size=384, grid=(5, 6), chip=CHIP, margin=16)


def per_pair(lib, ref, tgt, s, precision="fp32"):
    """The reference's calling pattern: one addReference and one addTarget per pair."""
    n = len(s.ref_begin)
    worker = lib.sequential(n, (CHIP, CHIP), (WINDOW, WINDOW), RM, RF, ZF, precision=precision)
    for k in range(n):
        r0 = tuple(int(v) for v in s.ref_begin[k])
        t0 = tuple(int(v) for v in s.tgt_begin[k])
        lib.addReference(worker, ref, k, r0, (r0[0] + CHIP, r0[1] + CHIP))
        lib.addTarget(worker, tgt, k, t0, (t0[0] + WINDOW, t0[1] + WINDOW))
    offsets = np.array(lib.adjust(worker), np.float32)
    prods = {k: np.frombuffer(v, np.float32).copy() for k, v in lib.products(worker).items()}
    return offsets, prods


def assert_same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    for k in ("offsets", "coarse_peak", "refined_peak", "snr"):
        np.testing.assert_array_equal(a[1][k].view(np.uint32), b[1][k].view(np.uint32), err_msg=k)


def test_slc_map_file_equals_uploaded_raster(lib, scene, tmp_path):
    ref_file, tgt_file = tmp_path / "reference.slc", tmp_path / "secondary.slc"
    scene.reference.astype(np.complex64).tofile(ref_file)
    scene.target.astype(np.complex64).tofile(tgt_file)
    lines, samples = scene.reference.shape
    mapped = per_pair(lib, lib.slc_map(str(ref_file), samples, lines),
                      lib.slc_map(filename=str(tgt_file), samples=samples, lines=lines), scene)
    uploaded = per_pair(lib, lib.raster(scene.reference, lines, samples),
                        lib.raster(scene.target, lines, samples), scene)
    assert_same(mapped, uploaded)
    # and the shift is the scene's
    np.testing.assert_allclose(mapped[0], np.broadcast_to(scene.truth, mapped[0].shape), atol=0.02)
    # the pyre-free SLC component, opened from the file and from its data
    by_path = dom.SLC(shape=(lines, samples), path=ref_file).open()
    by_data = dom.SLC(data=scene.target).open()
    assert_same(per_pair(lib, by_path.raster, by_data.raster, scene), uploaded)
    # slc_getitem reads the mapped file's pixels back
    k = 7 * samples + 11
    assert lib.slc_getitem(by_path.raster, k) == complex(scene.reference.reshape(-1)[k])


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_arrays_are_rasters(lib, scene, precision):
    """NumPy arrays go straight into addReference / addTarget / addPairs: each
    is uploaded at its first use by the worker and that copy is reused."""
    lines, samples = scene.reference.shape
    ref = np.ascontiguousarray(scene.reference, np.complex64)
    tgt = np.ascontiguousarray(scene.target, np.complex64)
    uploaded = per_pair(lib, lib.raster(ref, lines, samples), lib.raster(tgt, lines, samples),
                        scene, precision)
    before = sys.getrefcount(ref), sys.getrefcount(tgt)
    direct = per_pair(lib, ref, tgt, scene, precision)
    assert_same(direct, uploaded)
    # the worker is gone: it held each array once, however many tiles it read
    assert (sys.getrefcount(ref), sys.getrefcount(tgt)) == before
    n = len(scene.ref_begin)
    worker = lib.sequential(n, (CHIP, CHIP), (WINDOW, WINDOW), RM, RF, ZF, precision=precision)
    lib.addPairs(worker, ref, tgt, 0, scene.ref_begin.astype(np.int64).tobytes(),
                 scene.tgt_begin.astype(np.int64).tobytes())
    refs = sys.getrefcount(ref)
    for k in range(n):
        r0 = tuple(int(v) for v in scene.ref_begin[k])
        lib.addReference(worker, ref, k, r0, (r0[0] + CHIP, r0[1] + CHIP))
    assert sys.getrefcount(ref) == refs
    np.testing.assert_array_equal(np.array(lib.adjust(worker), np.float32), uploaded[0])


def test_float32_arrays_are_amplitude_rasters(lib, scene):
    amp_r = np.abs(scene.reference).astype(np.float32)
    amp_t = np.abs(scene.target).astype(np.float32)
    lines, samples = amp_r.shape
    uploaded = per_pair(lib, lib.raster(amp_r, lines, samples, dtype="float32"),
                        lib.raster(amp_t, lines, samples, dtype="float32"), scene)
    assert_same(per_pair(lib, amp_r, amp_t, scene), uploaded)


def test_cpu_slc_capsule_is_refused_with_the_replacement(lib, scene):
    """The CPU module's raster ("ampcor::slc::capsule_t") maps the file through
    pyre types; the HIP worker names what to pass instead."""
    new = ctypes.pythonapi.PyCapsule_New
    new.restype = ctypes.py_object
    new.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    name = ctypes.create_string_buffer(b"ampcor::slc::capsule_t")
    cpu_slc = new(ctypes.addressof(name), name, None)
    worker = lib.sequential(1, (CHIP, CHIP), (WINDOW, WINDOW), RM, RF, ZF)
    with pytest.raises(TypeError, match="slc_map"):
        lib.addReference(worker, cpu_slc, 0, (0, 0), (CHIP, CHIP))
    with pytest.raises(TypeError, match="slc_map"):
        lib.addPairs(worker, cpu_slc, cpu_slc, 0, np.zeros(2, np.int64).tobytes(),
                     np.zeros(2, np.int64).tobytes())


@pytest.mark.parametrize("bad", [
    lambda x: x.astype(np.complex128),      # not complex64 / float32
    lambda x: x.reshape(-1),                # not 2-D
    lambda x: x[:, ::2],                    # not contiguous
    lambda x: x.astype(np.int32),
])
def test_malformed_arrays_are_refused(lib, scene, bad):
    worker = lib.sequential(1, (CHIP, CHIP), (WINDOW, WINDOW), RM, RF, ZF)
    arr = bad(np.ascontiguousarray(scene.reference.real, np.float32))
    with pytest.raises((TypeError, ValueError, BufferError)):
        lib.addReference(worker, arr, 0, (0, 0), (CHIP, CHIP))
