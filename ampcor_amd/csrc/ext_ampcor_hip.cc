// ext_ampcor_hip.cc -- the CPython module "ampcor_hip", a thin layer over the C ABI.
//
// Same entry points, argument formats and results as the CUDA module
// (ext/ampcor_cuda/ampcor_cuda.cc:23-44, ext/ampcor_cuda/sequential.cc):
//   sequential(pairs, (refLines, refSamples), (tgtLines, tgtSamples),
//              refineMargin, refineFactor, zoomFactor[, precision, device]) -> worker
//   addReference(worker, raster, idx, (bLine, bSample), (eLine, eSample))  -> None
//   addTarget(worker, raster, idx, (bLine, bSample), (eLine, eSample))     -> None
//     (raster: a device raster capsule, or a (lines, samples) complex64 /
//     float32 array, uploaded once per worker)
//   adjust(worker) -> tuple of (dLine, dSample), one per pair
// plus the raster constructors that replace the host SLC capsule of the CPU
// module (ext/ampcor/slc.cc:50-140): slc_map(filename, samples, lines),
// raster(buffer, lines, samples[, dtype, device]), slc_pixelSize(),
// slc_getitem(raster, offset); the CPU module's refine(worker); and the
// batched and product accessors: addPairs, products, dims, stage_gamma,
// stage_gamma_refined, stage_locations, stage_upsample, stage_contract, stage_window_sums,
// stage_amplitudes, stage_zoom, zoom_stats,
// fp8_guard, set_timing, stage_times.
//
// Unlike the reference, adjust releases the GIL (a worker busy in another
// thread raises RuntimeError rather than racing it), and failures raise
// ampcor_hip.Error instead of escaping as C++ exceptions.
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "ampcor_hip.h"

namespace {

const char *const kWorkerCapsule = "ampcor::hip::sequential::capsule_t";
const char *const kRasterCapsule = "ampcor::hip::raster::capsule_t";
const char *const kServerCapsule = "ampcor::hip::server::capsule_t";

PyObject *g_error = nullptr;

struct Worker {
    ampcor_hip_plan *plan;
    PyObject *rasters;  // list: keeps every raster the plan points into alive
    size_t pairs;
    bool busy;          // a call released the GIL inside the plan (adjust, autotune)
    int device;         // the plan's device
    // buffers (NumPy arrays) passed as rasters: the device copy made at their
    // first use, by object (the object itself is kept alive in `rasters`)
    std::unordered_map<PyObject *, ampcor_hip_raster *> uploads;
};

PyObject *raise(int status)
{
    PyErr_Format(status == AMPCOR_HIP_EINVAL ? PyExc_ValueError : g_error, "%s (status %d)",
                 ampcor_hip_last_error(), status);
    return nullptr;
}

void free_worker(PyObject *capsule)
{
    if (!PyCapsule_IsValid(capsule, kWorkerCapsule)) return;
    auto *w = static_cast<Worker *>(PyCapsule_GetPointer(capsule, kWorkerCapsule));
    ampcor_hip_plan_destroy(w->plan);
    Py_XDECREF(w->rasters);
    delete w;
}

void free_raster(PyObject *capsule)
{
    if (!PyCapsule_IsValid(capsule, kRasterCapsule)) return;
    ampcor_hip_raster_destroy(static_cast<ampcor_hip_raster *>(PyCapsule_GetPointer(capsule, kRasterCapsule)));
}

// the worker of a capsule; a worker that another thread is using (inside
// adjust or autotune, with the GIL released) is refused rather than raced
Worker *get_worker(PyObject *o)
{
    if (!PyCapsule_IsValid(o, kWorkerCapsule)) {
        PyErr_SetString(PyExc_TypeError, "invalid HIP Sequential worker capsule");
        return nullptr;
    }
    auto *w = static_cast<Worker *>(PyCapsule_GetPointer(o, kWorkerCapsule));
    if (w->busy) {
        PyErr_SetString(PyExc_RuntimeError, "the worker is in use by another thread");
        return nullptr;
    }
    return w;
}

ampcor_hip_raster *get_raster(PyObject *o)
{
    if (!PyCapsule_IsValid(o, kRasterCapsule)) {
        PyErr_SetString(PyExc_TypeError, "invalid raster capsule");
        return nullptr;
    }
    return static_cast<ampcor_hip_raster *>(PyCapsule_GetPointer(o, kRasterCapsule));
}

// the CPU module's SLC capsule (ext/ampcor/capsules.h:26, made by slc_map in
// ext/ampcor/slc.cc:50-93): a memory map inside that module, read through its
// own C++ types, which this module does not link
const char *const kHostSlcCapsule = "ampcor::slc::capsule_t";

// keep `o` alive as long as the worker (by identity: `==` on arrays compares pixels)
int worker_keeps(Worker *w, PyObject *o)
{
    const Py_ssize_t n = PyList_GET_SIZE(w->rasters);
    for (Py_ssize_t k = 0; k < n; ++k)
        if (PyList_GET_ITEM(w->rasters, k) == o) return 0;
    return PyList_Append(w->rasters, o);
}

// the device raster behind a raster argument of addReference / addTarget /
// addPairs (ext/ampcor_cuda/sequential.cc:85-143 takes the SLC capsule there):
//   - a raster capsule of this module (raster(), slc_map()): used as it is;
//   - a 2-D C-contiguous buffer of complex64 or float32 pixels (a NumPy array
//     of shape (lines, samples)): uploaded at its first use by this worker, and
//     that copy serves every later tile of the worker (the pixels are read
//     once: changes to the array after its first use are not seen);
//   - the CPU module's SLC capsule: TypeError naming the replacements.
// The worker keeps the argument alive.
ampcor_hip_raster *worker_raster(Worker *w, PyObject *o)
{
    if (PyCapsule_CheckExact(o)) {
        const char *name = PyCapsule_GetName(o);
        if (name && !strcmp(name, kRasterCapsule)) {
            if (worker_keeps(w, o)) return nullptr;
            return static_cast<ampcor_hip_raster *>(PyCapsule_GetPointer(o, kRasterCapsule));
        }
        if (name && !strcmp(name, kHostSlcCapsule)) {
            PyErr_SetString(PyExc_TypeError,
                            "an SLC capsule of the CPU module (ampcor.ext.ampcor.slc_map) maps "
                            "the raster into host memory through types this module does not "
                            "link; open the same file with ampcor_hip.slc_map(filename, samples, "
                            "lines), or pass the pixels as a (lines, samples) complex64 array");
            return nullptr;
        }
        PyErr_SetString(PyExc_TypeError, "invalid raster capsule");
        return nullptr;
    }
    auto hit = w->uploads.find(o);
    if (hit != w->uploads.end()) return hit->second;
    if (!PyObject_CheckBuffer(o)) {
        PyErr_SetString(PyExc_TypeError, "a raster is a raster capsule (ampcor_hip.raster, "
                                         "ampcor_hip.slc_map) or a 2-D complex64 / float32 array");
        return nullptr;
    }
    Py_buffer view;
    if (PyObject_GetBuffer(o, &view, PyBUF_C_CONTIGUOUS | PyBUF_FORMAT)) return nullptr;
    // the pixel type from the buffer's struct format, native byte order
    const char *fmt = view.format ? view.format : "B";
    if (*fmt == '@' || *fmt == '=' || *fmt == '<') ++fmt;
    int dtype = -1;
    if (!strcmp(fmt, "Zf") && view.itemsize == 8) dtype = AMPCOR_HIP_COMPLEX64;
    else if (!strcmp(fmt, "f") && view.itemsize == 4) dtype = AMPCOR_HIP_FLOAT32;
    if (dtype < 0 || view.ndim != 2) {
        PyErr_Format(PyExc_TypeError,
                     "a raster array must be 2-D (lines, samples) complex64 or float32 in native "
                     "byte order (got %d-D, format '%s')",
                     view.ndim, view.format ? view.format : "B");
        PyBuffer_Release(&view);
        return nullptr;
    }
    const size_t lines = (size_t)view.shape[0], samples = (size_t)view.shape[1];
    ampcor_hip_raster *r = nullptr;
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_raster_create(w->device, dtype, view.buf, lines, samples, &r);
    Py_END_ALLOW_THREADS
    PyBuffer_Release(&view);
    if (rc) {
        raise(rc);
        return nullptr;
    }
    PyObject *capsule = PyCapsule_New(r, kRasterCapsule, free_raster);
    if (!capsule) {
        ampcor_hip_raster_destroy(r);
        return nullptr;
    }
    const int kept = worker_keeps(w, capsule) || worker_keeps(w, o);
    Py_DECREF(capsule);   // the worker's list holds it
    if (kept) return nullptr;
    w->uploads[o] = r;
    return r;
}

int parse_dtype(const char *name, int *dtype)
{
    if (!strcmp(name, "complex64")) *dtype = AMPCOR_HIP_COMPLEX64;
    else if (!strcmp(name, "float32")) *dtype = AMPCOR_HIP_FLOAT32;
    else {
        PyErr_Format(PyExc_ValueError, "unknown raster pixel type '%s'", name);
        return -1;
    }
    return 0;
}

int parse_precision(const char *name, int *precision)
{
    if (!strcmp(name, "fp32")) *precision = AMPCOR_HIP_FP32;
    else if (!strcmp(name, "bf16")) *precision = AMPCOR_HIP_BF16;
    else if (!strcmp(name, "fp16")) *precision = AMPCOR_HIP_FP16;
    else if (!strcmp(name, "bf16x3")) *precision = AMPCOR_HIP_BF16X3;
    else if (!strcmp(name, "fp16x3")) *precision = AMPCOR_HIP_FP16X3;
    else if (!strcmp(name, "fp8")) *precision = AMPCOR_HIP_FP8;
    else {
        PyErr_Format(PyExc_ValueError, "unknown precision '%s'", name);
        return -1;
    }
    return 0;
}

// ---- rasters -----------------------------------------------------------------

PyObject *slc_pixelSize(PyObject *, PyObject *) { return PyLong_FromLong(8); }

// slc_getitem(raster, offset) -> complex: the pixel at row-major offset
// (ext/ampcor/slc.cc:105-140); a float32 raster's pixels have no imaginary part.
// Past the end raises IndexError (the reference reads past its map).
PyObject *slc_getitem(PyObject *, PyObject *args)
{
    PyObject *capsule;
    unsigned long long offset;
    if (!PyArg_ParseTuple(args, "OK:slc_getitem", &capsule, &offset)) return nullptr;
    ampcor_hip_raster *r = get_raster(capsule);
    if (!r) return nullptr;
    size_t lines = 0, samples = 0;
    if (int rc = ampcor_hip_raster_shape(r, &lines, &samples)) return raise(rc);
    if (offset >= (unsigned long long)(lines * samples)) {
        PyErr_SetString(PyExc_IndexError, "raster offset out of range");
        return nullptr;
    }
    float px[2] = {0.0f, 0.0f};
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_raster_read(r, (size_t)offset, 1, px);
    Py_END_ALLOW_THREADS
    if (rc) return raise(rc);
    return PyComplex_FromDoubles(px[0], px[1]);
}

PyObject *raster(PyObject *, PyObject *args, PyObject *kwds)
{
    Py_buffer view;
    unsigned long lines = 0, samples = 0;
    const char *dtype_name = "complex64";
    int device = 0;
    static const char *names[] = {"data", "lines", "samples", "dtype", "device", nullptr};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "y*kk|si:raster", const_cast<char **>(names),
                                     &view, &lines, &samples, &dtype_name, &device))
        return nullptr;
    int dtype;
    if (parse_dtype(dtype_name, &dtype)) {
        PyBuffer_Release(&view);
        return nullptr;
    }
    const size_t need = (size_t)lines * samples * (dtype == AMPCOR_HIP_COMPLEX64 ? 8 : 4);
    if ((size_t)view.len != need) {
        PyBuffer_Release(&view);
        PyErr_Format(PyExc_ValueError, "raster buffer holds %zd bytes, shape needs %zu", view.len, need);
        return nullptr;
    }
    ampcor_hip_raster *r = nullptr;
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_raster_create(device, dtype, view.buf, lines, samples, &r);
    Py_END_ALLOW_THREADS
    PyBuffer_Release(&view);
    if (rc) return raise(rc);
    return PyCapsule_New(r, kRasterCapsule, free_raster);
}

// memory-map a raw complex64 raster file and upload it (ext/ampcor/slc.cc:50-93)
PyObject *slc_map(PyObject *, PyObject *args, PyObject *kwds)
{
    const char *filename;
    unsigned long samples = 0, lines = 0;
    int device = 0;
    static const char *names[] = {"filename", "samples", "lines", "device", nullptr};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "skk|i:slc_map", const_cast<char **>(names),
                                     &filename, &samples, &lines, &device))
        return nullptr;
    const size_t bytes = (size_t)lines * samples * 8;
    int fd = open(filename, O_RDONLY);
    if (fd < 0) return PyErr_SetFromErrnoWithFilename(PyExc_OSError, filename);
    struct stat st;
    if (fstat(fd, &st) != 0 || (size_t)st.st_size < bytes) {
        close(fd);
        PyErr_Format(PyExc_ValueError, "'%s' is smaller than %lu x %lu complex64 pixels", filename,
                     lines, samples);
        return nullptr;
    }
    void *map = bytes ? mmap(nullptr, bytes, PROT_READ, MAP_PRIVATE, fd, 0) : MAP_FAILED;
    close(fd);
    if (map == MAP_FAILED) return PyErr_SetFromErrnoWithFilename(PyExc_OSError, filename);
    ampcor_hip_raster *r = nullptr;
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_raster_create(device, AMPCOR_HIP_COMPLEX64, map, lines, samples, &r);
    Py_END_ALLOW_THREADS
    munmap(map, bytes);
    if (rc) return raise(rc);
    return PyCapsule_New(r, kRasterCapsule, free_raster);
}

// ---- the worker ----------------------------------------------------------------

PyObject *sequential(PyObject *, PyObject *args, PyObject *kwds)
{
    unsigned long pairs, refLines, refSamples, tgtLines, tgtSamples;
    unsigned long refineMargin, refineFactor, zoomFactor;
    const char *precision_name = "fp32";
    int device = 0;
    static const char *names[] = {"pairs", "chip", "window", "refineMargin", "refineFactor",
                                  "zoomFactor", "precision", "device", nullptr};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "k(kk)(kk)kkk|si:sequential",
                                     const_cast<char **>(names), &pairs, &refLines, &refSamples,
                                     &tgtLines, &tgtSamples, &refineMargin, &refineFactor,
                                     &zoomFactor, &precision_name, &device))
        return nullptr;
    int precision;
    if (parse_precision(precision_name, &precision)) return nullptr;
    ampcor_hip_plan *plan = nullptr;
    int rc = ampcor_hip_plan_create(device, pairs, refLines, refSamples, tgtLines, tgtSamples,
                                    refineMargin, refineFactor, zoomFactor, precision, &plan);
    if (rc) return raise(rc);
    auto *w = new Worker{plan, PyList_New(0), pairs, false, device, {}};
    return PyCapsule_New(w, kWorkerCapsule, free_worker);
}

PyObject *add_tile(PyObject *args, bool reference)
{
    PyObject *pyWorker, *pyRaster;
    unsigned long idx, bl, bs, el, es;
    if (!PyArg_ParseTuple(args, reference ? "O!Ok(kk)(kk):addReference" : "O!Ok(kk)(kk):addTarget",
                          &PyCapsule_Type, &pyWorker, &pyRaster, &idx, &bl, &bs, &el, &es))
        return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    ampcor_hip_raster *r = worker_raster(w, pyRaster);
    if (!r) return nullptr;
    int rc = reference ? ampcor_hip_add_reference(w->plan, r, idx, bl, bs, el, es)
                       : ampcor_hip_add_target(w->plan, r, idx, bl, bs, el, es);
    if (rc) return raise(rc);
    Py_RETURN_NONE;
}

PyObject *addReference(PyObject *, PyObject *args) { return add_tile(args, true); }
PyObject *addTarget(PyObject *, PyObject *args) { return add_tile(args, false); }

// addPairs(worker, refRaster, tgtRaster, first, refBegins, tgtBegins): the two
// corner arrays are buffers of int64 (line, sample) pairs
PyObject *addPairs(PyObject *, PyObject *args)
{
    PyObject *pyWorker, *pyRef, *pyTgt;
    unsigned long first;
    Py_buffer rb, tb;
    if (!PyArg_ParseTuple(args, "O!OOky*y*:addPairs", &PyCapsule_Type, &pyWorker, &pyRef, &pyTgt,
                          &first, &rb, &tb))
        return nullptr;
    PyObject *result = nullptr;
    Worker *w = get_worker(pyWorker);
    ampcor_hip_raster *ref = w ? worker_raster(w, pyRef) : nullptr;
    ampcor_hip_raster *tgt = ref ? worker_raster(w, pyTgt) : nullptr;
    if (tgt) {
        if (rb.len != tb.len || rb.len % 16) {
            PyErr_SetString(PyExc_ValueError, "corner buffers must hold equal counts of int64 pairs");
        } else {
            int rc = ampcor_hip_add_pairs(w->plan, ref, tgt, first, rb.len / 16,
                                          static_cast<const int64_t *>(rb.buf),
                                          static_cast<const int64_t *>(tb.buf));
            if (rc) raise(rc);
            else {
                Py_INCREF(Py_None);
                result = Py_None;
            }
        }
    }
    PyBuffer_Release(&rb);
    PyBuffer_Release(&tb);
    return result;
}

PyObject *adjust(PyObject *, PyObject *args, PyObject *kwds)
{
    // offsets=False skips building the tuple of (dy, dx) pairs: a caller that
    // reads the products afterwards (products()) does not pay for the objects
    static const char *kw[] = {"worker", "offsets", nullptr};
    PyObject *pyWorker;
    int want = 1;
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "O!|p:adjust", const_cast<char **>(kw),
                                     &PyCapsule_Type, &pyWorker, &want))
        return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    std::vector<float> offsets(want ? 2 * w->pairs : 0);
    int rc;
    w->busy = true;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_adjust(w->plan, want ? offsets.data() : nullptr);
    Py_END_ALLOW_THREADS
    w->busy = false;
    if (rc) return raise(rc);
    if (!want) Py_RETURN_NONE;
    PyObject *results = PyTuple_New(w->pairs);
    if (!results) return nullptr;
    for (size_t p = 0; p < w->pairs; ++p) {
        PyObject *shift = PyTuple_New(2);
        if (!shift) {
            Py_DECREF(results);
            return nullptr;
        }
        PyTuple_SET_ITEM(shift, 0, PyFloat_FromDouble(offsets[2 * p]));
        PyTuple_SET_ITEM(shift, 1, PyFloat_FromDouble(offsets[2 * p + 1]));
        PyTuple_SET_ITEM(results, p, shift);
    }
    return results;
}

// refine(worker) -> None: the CPU module's separate refinement call
// (ext/ampcor/sequential.cc:262-294, whose worker.refine() is empty,
// lib/libampcor/correlators/Sequential.cc:101-104).  adjust already runs the
// whole refinement chain, so this only checks the worker.
PyObject *refine(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    if (!PyArg_ParseTuple(args, "O!:refine", &PyCapsule_Type, &pyWorker)) return nullptr;
    if (!get_worker(pyWorker)) return nullptr;
    Py_RETURN_NONE;
}

PyObject *bytes_of(size_t n)
{
    return PyBytes_FromStringAndSize(nullptr, (Py_ssize_t)n);
}

PyObject *products(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    if (!PyArg_ParseTuple(args, "O!:products", &PyCapsule_Type, &pyWorker)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    const size_t N = w->pairs;
    PyObject *off = bytes_of(8 * N), *cp = bytes_of(4 * N), *rp = bytes_of(4 * N), *sn = bytes_of(4 * N);
    if (!off || !cp || !rp || !sn) {
        Py_XDECREF(off); Py_XDECREF(cp); Py_XDECREF(rp); Py_XDECREF(sn);
        return nullptr;
    }
    int rc = ampcor_hip_products(w->plan, reinterpret_cast<float *>(PyBytes_AS_STRING(off)),
                                 reinterpret_cast<float *>(PyBytes_AS_STRING(cp)),
                                 reinterpret_cast<float *>(PyBytes_AS_STRING(rp)),
                                 reinterpret_cast<float *>(PyBytes_AS_STRING(sn)));
    if (rc) {
        Py_DECREF(off); Py_DECREF(cp); Py_DECREF(rp); Py_DECREF(sn);
        return raise(rc);
    }
    return Py_BuildValue("{sNsNsNsN}", "offsets", off, "coarse_peak", cp, "refined_peak", rp, "snr", sn);
}

PyObject *dims(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    if (!PyArg_ParseTuple(args, "O!:dims", &PyCapsule_Type, &pyWorker)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    int P, R, Z;
    if (int rc = ampcor_hip_plan_dims(w->plan, &P, &R, &Z)) return raise(rc);
    return Py_BuildValue("(iii)", P, R, Z);
}

PyObject *stage_gamma_impl(PyObject *args, bool refined)
{
    PyObject *pyWorker;
    if (!PyArg_ParseTuple(args, "O!", &PyCapsule_Type, &pyWorker)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    int P, R, Z;
    if (int rc = ampcor_hip_plan_dims(w->plan, &P, &R, &Z)) return raise(rc);
    const size_t d = refined ? R : P;
    PyObject *out = bytes_of(4 * w->pairs * d * d);
    if (!out) return nullptr;
    float *dst = reinterpret_cast<float *>(PyBytes_AS_STRING(out));
    int rc = refined ? ampcor_hip_stage_gamma_refined(w->plan, dst) : ampcor_hip_stage_gamma(w->plan, dst);
    if (rc) {
        Py_DECREF(out);
        return raise(rc);
    }
    return out;
}

PyObject *stage_gamma(PyObject *, PyObject *args) { return stage_gamma_impl(args, false); }
PyObject *stage_gamma_refined(PyObject *, PyObject *args) { return stage_gamma_impl(args, true); }

PyObject *stage_locations(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    if (!PyArg_ParseTuple(args, "O!:stage_locations", &PyCapsule_Type, &pyWorker)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    const size_t n = 8 * w->pairs;
    PyObject *c = bytes_of(n), *wi = bytes_of(n), *z = bytes_of(n);
    if (!c || !wi || !z) {
        Py_XDECREF(c); Py_XDECREF(wi); Py_XDECREF(z);
        return nullptr;
    }
    int rc = ampcor_hip_stage_locations(w->plan, reinterpret_cast<int32_t *>(PyBytes_AS_STRING(c)),
                                        reinterpret_cast<int32_t *>(PyBytes_AS_STRING(wi)),
                                        reinterpret_cast<int32_t *>(PyBytes_AS_STRING(z)));
    if (rc) {
        Py_DECREF(c); Py_DECREF(wi); Py_DECREF(z);
        return raise(rc);
    }
    return Py_BuildValue("{sNsNsN}", "coarse", c, "window", wi, "zoomed", z);
}

// stage_upsample(worker, tiles, count, dtype="complex64") -> bytes: the target
// upsampling of the worker's form on count tiles [count][E][E]; returns |y|
// float32 [count][rf E][rf E]
PyObject *stage_upsample(PyObject *, PyObject *args, PyObject *kwds)
{
    PyObject *pyWorker;
    Py_buffer view;
    unsigned long count;
    const char *dtype_name = "complex64";
    static const char *names[] = {"worker", "tiles", "count", "dtype", nullptr};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "O!y*k|s:stage_upsample",
                                     const_cast<char **>(names), &PyCapsule_Type, &pyWorker, &view,
                                     &count, &dtype_name))
        return nullptr;
    int dtype = 0;
    Worker *w = get_worker(pyWorker);
    if (!w || parse_dtype(dtype_name, &dtype)) {
        PyBuffer_Release(&view);
        return nullptr;
    }
    int P, R, Z;
    if (int rc = ampcor_hip_plan_dims(w->plan, &P, &R, &Z)) {
        PyBuffer_Release(&view);
        return raise(rc);
    }
    // E and rf from the plan's dimensions: R = 2 rf rm + 1, rf E = refined window
    size_t E = 0, M = 0;
    if (int rc = ampcor_hip_plan_tile_dims(w->plan, &E, &M)) {
        PyBuffer_Release(&view);
        return raise(rc);
    }
    const size_t px = dtype == AMPCOR_HIP_COMPLEX64 ? 8 : 4;
    if ((size_t)view.len != count * E * E * px) {
        PyBuffer_Release(&view);
        PyErr_Format(PyExc_ValueError, "tiles hold %zd bytes, %lu tiles of %zu x %zu need %zu",
                     view.len, count, E, E, count * E * E * px);
        return nullptr;
    }
    PyObject *out = bytes_of(4 * count * M * M);
    if (!out) {
        PyBuffer_Release(&view);
        return nullptr;
    }
    int rc;
    w->busy = true;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_stage_upsample(w->plan, dtype, view.buf, count,
                                   reinterpret_cast<float *>(PyBytes_AS_STRING(out)));
    Py_END_ALLOW_THREADS
    w->busy = false;
    PyBuffer_Release(&view);
    if (rc) {
        Py_DECREF(out);
        return raise(rc);
    }
    return out;
}

// stage_contract(worker, stage, form, chip, c, window, w, count) -> bytes: one
// contraction of a plain BF16 / FP16 worker on 16-bit operand patterns
// chip [count][c][c], window [count][w][w]; returns float32 numerators
// [count][p][p], p = w - c + 1 (form: 0 current, 1 tiles, 2 streaming, 3 packed,
// 4 / 5 packed in two / four chip-row ranges)
PyObject *stage_contract(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    int stage, form, c, wd;
    Py_buffer chip, win;
    unsigned long count;
    if (!PyArg_ParseTuple(args, "O!iiy*iy*ik:stage_contract", &PyCapsule_Type, &pyWorker, &stage,
                          &form, &chip, &c, &win, &wd, &count))
        return nullptr;
    Worker *w = get_worker(pyWorker);
    auto release = [&] {
        PyBuffer_Release(&chip);
        PyBuffer_Release(&win);
    };
    if (!w) {
        release();
        return nullptr;
    }
    if (c <= 0 || wd < c || (size_t)chip.len != 2 * count * (size_t)c * c ||
        (size_t)win.len != 2 * count * (size_t)wd * wd) {
        release();
        PyErr_SetString(PyExc_ValueError, "chip / window buffers do not hold count x c^2 / w^2 "
                                          "16-bit values");
        return nullptr;
    }
    const size_t p = (size_t)(wd - c + 1);
    PyObject *out = bytes_of(4 * count * p * p);
    if (!out) {
        release();
        return nullptr;
    }
    int rc;
    w->busy = true;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_stage_contract(w->plan, stage, form, static_cast<const uint16_t *>(chip.buf), c,
                                   static_cast<const uint16_t *>(win.buf), wd, count,
                                   reinterpret_cast<float *>(PyBytes_AS_STRING(out)));
    Py_END_ALLOW_THREADS
    w->busy = false;
    release();
    if (rc) {
        Py_DECREF(out);
        return raise(rc);
    }
    return out;
}

// stage_window_sums(worker, stage, windows, w, count) -> bytes: the FP64 window
// sums (sum t, sum t^2) over every chip-sized placement of float32 windows
// [count][w][w]; returns float64 [count][p][p][2]
PyObject *stage_window_sums(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    int stage, wd;
    Py_buffer win;
    unsigned long count;
    if (!PyArg_ParseTuple(args, "O!iy*ik:stage_window_sums", &PyCapsule_Type, &pyWorker, &stage,
                          &win, &wd, &count))
        return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) {
        PyBuffer_Release(&win);
        return nullptr;
    }
    int P, R, Z;
    int rc = ampcor_hip_plan_dims(w->plan, &P, &R, &Z);
    if (rc) {
        PyBuffer_Release(&win);
        return raise(rc);
    }
    if (wd <= 0 || (size_t)win.len != 4 * count * (size_t)wd * wd || (stage != 0 && stage != 1)) {
        PyBuffer_Release(&win);
        PyErr_SetString(PyExc_ValueError, "windows do not hold count x w^2 float32 values, or "
                                          "stage is not 0 or 1");
        return nullptr;
    }
    const size_t p = (size_t)(stage ? R : P);
    PyObject *out = bytes_of(16 * count * p * p);
    if (!out) {
        PyBuffer_Release(&win);
        return nullptr;
    }
    w->busy = true;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_stage_window_sums(w->plan, stage, static_cast<const float *>(win.buf), wd, count,
                                      reinterpret_cast<double *>(PyBytes_AS_STRING(out)));
    Py_END_ALLOW_THREADS
    w->busy = false;
    PyBuffer_Release(&win);
    if (rc) {
        Py_DECREF(out);
        return raise(rc);
    }
    return out;
}

// stage_amplitudes(worker, stage, chip, window) -> (chips, windows): float32
// bytes of the amplitudes the last adjust of an FP32 worker worked on, stage 0
// the detected ones, 1 the refined ones: [pairs][chip][chip] and
// [pairs][window][window]
PyObject *stage_amplitudes(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    int stage, c, wd;
    if (!PyArg_ParseTuple(args, "O!iii:stage_amplitudes", &PyCapsule_Type, &pyWorker, &stage, &c,
                          &wd))
        return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    if (c <= 0 || wd < c) {
        PyErr_SetString(PyExc_ValueError, "chip and window must be positive, window >= chip");
        return nullptr;
    }
    const size_t pairs = w->pairs, C = (size_t)c, W = (size_t)wd;
    int rc;
    PyObject *chips = bytes_of(4 * pairs * C * C);
    PyObject *wins = chips ? bytes_of(4 * pairs * W * W) : nullptr;
    if (!chips || !wins) {
        Py_XDECREF(chips);
        return nullptr;
    }
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_stage_amplitudes(w->plan, stage, c, wd,
                                     reinterpret_cast<float *>(PyBytes_AS_STRING(chips)),
                                     reinterpret_cast<float *>(PyBytes_AS_STRING(wins)));
    Py_END_ALLOW_THREADS
    if (rc) {
        Py_DECREF(chips);
        Py_DECREF(wins);
        return raise(rc);
    }
    return Py_BuildValue("(NN)", chips, wins);
}

// stage_zoom(worker, surfaces): the zoom stage alone on float32 refined
// surfaces [pairs][R][R]; products() and stage_locations() then read its results
PyObject *stage_zoom(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    Py_buffer view;
    if (!PyArg_ParseTuple(args, "O!y*:stage_zoom", &PyCapsule_Type, &pyWorker, &view)) return nullptr;
    Worker *w = get_worker(pyWorker);
    int P, R, Z;
    int rc = w ? ampcor_hip_plan_dims(w->plan, &P, &R, &Z) : 0;
    if (!w || rc) {
        PyBuffer_Release(&view);
        return w ? raise(rc) : nullptr;
    }
    if ((size_t)view.len != 4 * w->pairs * (size_t)R * R) {
        PyBuffer_Release(&view);
        PyErr_Format(PyExc_ValueError, "surfaces hold %zd bytes, the plan needs %zu", view.len,
                     4 * w->pairs * (size_t)R * R);
        return nullptr;
    }
    w->busy = true;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_stage_zoom(w->plan, static_cast<const float *>(view.buf));
    Py_END_ALLOW_THREADS
    w->busy = false;
    PyBuffer_Release(&view);
    if (rc) return raise(rc);
    Py_RETURN_NONE;
}

// zoom_stats(worker) -> {"cells", "pairs", "searched"}: the zoom's whole-grid searches
PyObject *zoom_stats(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    if (!PyArg_ParseTuple(args, "O!:zoom_stats", &PyCapsule_Type, &pyWorker)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    int64_t cells = 0, pairs = 0, searched = 0;
    if (int rc = ampcor_hip_zoom_stats(w->plan, &cells, &pairs, &searched)) return raise(rc);
    return Py_BuildValue("{sLsLsL}", "cells", (long long)cells, "pairs", (long long)pairs,
                         "searched", (long long)searched);
}

// fp8_guard(worker) -> {"rescored", "moved", "overflow"}: the FP8 form's peak
// guard in the last adjust (zeros for the other forms)
PyObject *fp8_guard(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    if (!PyArg_ParseTuple(args, "O!:fp8_guard", &PyCapsule_Type, &pyWorker)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    int64_t rescored = 0, moved = 0, overflow = 0;
    if (int rc = ampcor_hip_fp8_guard_stats(w->plan, &rescored, &moved, &overflow)) return raise(rc);
    return Py_BuildValue("{sLsLsL}", "rescored", (long long)rescored, "moved", (long long)moved,
                         "overflow", (long long)overflow);
}

PyObject *set_timing(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    int flag;
    if (!PyArg_ParseTuple(args, "O!p:set_timing", &PyCapsule_Type, &pyWorker, &flag)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    if (int rc = ampcor_hip_set_timing(w->plan, flag)) return raise(rc);
    Py_RETURN_NONE;
}

PyObject *stage_times(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    if (!PyArg_ParseTuple(args, "O!:stage_times", &PyCapsule_Type, &pyWorker)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    float ms[64];
    const char *names[64];
    int count = 0;
    if (int rc = ampcor_hip_stage_times(w->plan, 64, &count, ms, names)) return raise(rc);
    PyObject *d = PyDict_New();
    for (int k = 0; k < count && k < 64; ++k) {
        PyObject *v = PyFloat_FromDouble(ms[k]);
        PyDict_SetItemString(d, names[k], v);
        Py_DECREF(v);
    }
    return d;
}


// ---- the serving loop ------------------------------------------------------------

struct Server {
    ampcor_hip_server *server;
    // every job holds its two rasters alive until it is waited for
    std::unordered_map<uint64_t, std::pair<PyObject *, PyObject *>> held;
};

void free_server(PyObject *capsule)
{
    if (!PyCapsule_IsValid(capsule, kServerCapsule)) return;
    auto *s = static_cast<Server *>(PyCapsule_GetPointer(capsule, kServerCapsule));
    Py_BEGIN_ALLOW_THREADS
    ampcor_hip_server_destroy(s->server);
    Py_END_ALLOW_THREADS
    for (auto &kv : s->held) {
        Py_XDECREF(kv.second.first);
        Py_XDECREF(kv.second.second);
    }
    delete s;
}

// drop the references a job held on its rasters
void release_held(Server *s, uint64_t id)
{
    auto it = s->held.find(id);
    if (it == s->held.end()) return;
    Py_XDECREF(it->second.first);
    Py_XDECREF(it->second.second);
    s->held.erase(it);
}

Server *get_server(PyObject *o)
{
    if (!PyCapsule_IsValid(o, kServerCapsule)) {
        PyErr_SetString(PyExc_TypeError, "invalid server capsule");
        return nullptr;
    }
    return static_cast<Server *>(PyCapsule_GetPointer(o, kServerCapsule));
}

PyObject *server(PyObject *, PyObject *args, PyObject *kwds)
{
    int device = 0, slots = 2, graphs = 1;
    double budget_gb = 64.0;
    unsigned long batch = 4096;
    static const char *names[] = {"device", "budget_gb", "batch", "slots", "graphs", nullptr};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "|idkip:server", const_cast<char **>(names), &device,
                                     &budget_gb, &batch, &slots, &graphs))
        return nullptr;
    ampcor_hip_server *sv = nullptr;
    int rc = ampcor_hip_server_create(device, (size_t)(budget_gb * (1ull << 30)), batch, slots, graphs, &sv);
    if (rc) return raise(rc);
    return PyCapsule_New(new Server{sv, {}}, kServerCapsule, free_server);
}

PyObject *submit(PyObject *, PyObject *args, PyObject *kwds)
{
    PyObject *pyServer, *pyRef, *pyTgt;
    Py_buffer rb, tb;
    unsigned long chip, window, rm, rf, zf;
    const char *precision_name = "fp32";
    static const char *names[] = {"server", "reference", "target", "refBegin", "tgtBegin", "chip",
                                  "window", "refineMargin", "refineFactor", "zoomFactor",
                                  "precision", nullptr};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "O!O!O!y*y*kkkkk|s:submit", const_cast<char **>(names),
                                     &PyCapsule_Type, &pyServer, &PyCapsule_Type, &pyRef,
                                     &PyCapsule_Type, &pyTgt, &rb, &tb, &chip, &window, &rm, &rf,
                                     &zf, &precision_name))
        return nullptr;
    PyObject *result = nullptr;
    Server *s = get_server(pyServer);
    ampcor_hip_raster *ref = s ? get_raster(pyRef) : nullptr;
    ampcor_hip_raster *tgt = ref ? get_raster(pyTgt) : nullptr;
    int precision;
    if (tgt && !parse_precision(precision_name, &precision)) {
        if (rb.len != tb.len || rb.len % 16 || rb.len == 0) {
            PyErr_SetString(PyExc_ValueError, "corner buffers must hold equal, nonzero counts of int64 pairs");
        } else {
            uint64_t id = 0;
            int rc = ampcor_hip_server_submit(s->server, ref, tgt, rb.len / 16,
                                              static_cast<const int64_t *>(rb.buf),
                                              static_cast<const int64_t *>(tb.buf), chip, window, rm,
                                              rf, zf, precision, &id);
            if (rc) raise(rc);
            else {
                Py_INCREF(pyRef);
                Py_INCREF(pyTgt);
                s->held[id] = {pyRef, pyTgt};
                result = PyLong_FromUnsignedLongLong(id);
            }
        }
    }
    PyBuffer_Release(&rb);
    PyBuffer_Release(&tb);
    return result;
}

PyObject *wait(PyObject *, PyObject *args, PyObject *kwds)
{
    PyObject *pyServer;
    unsigned long long id;
    int timeout_ms = -1;
    static const char *names[] = {"server", "job", "timeout_ms", nullptr};
    if (!PyArg_ParseTupleAndKeywords(args, kwds, "O!K|i:wait", const_cast<char **>(names),
                                     &PyCapsule_Type, &pyServer, &id, &timeout_ms))
        return nullptr;
    Server *s = get_server(pyServer);
    if (!s) return nullptr;
    size_t pairs = 0;
    if (int rc = ampcor_hip_server_job_pairs(s->server, id, &pairs)) return raise(rc);
    PyObject *off = bytes_of(8 * pairs), *cp = bytes_of(4 * pairs), *rp = bytes_of(4 * pairs),
             *sn = bytes_of(4 * pairs);
    if (!off || !cp || !rp || !sn) {
        Py_XDECREF(off); Py_XDECREF(cp); Py_XDECREF(rp); Py_XDECREF(sn);
        return nullptr;
    }
    double latency = 0.0;
    int rc;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_server_wait(s->server, id, timeout_ms, pairs,
                                reinterpret_cast<float *>(PyBytes_AS_STRING(off)),
                                reinterpret_cast<float *>(PyBytes_AS_STRING(cp)),
                                reinterpret_cast<float *>(PyBytes_AS_STRING(rp)),
                                reinterpret_cast<float *>(PyBytes_AS_STRING(sn)), &latency);
    Py_END_ALLOW_THREADS
    if (rc) {
        Py_DECREF(off); Py_DECREF(cp); Py_DECREF(rp); Py_DECREF(sn);
        // a job that failed alone is forgotten by the server: its rasters may go
        const std::string text = ampcor_hip_last_error();
        size_t left = 0;
        if (ampcor_hip_server_job_pairs(s->server, id, &left) != AMPCOR_HIP_OK) release_held(s, id);
        PyErr_Format(rc == AMPCOR_HIP_EINVAL ? PyExc_ValueError : g_error, "%s (status %d)",
                     text.c_str(), rc);
        return nullptr;
    }
    release_held(s, id);
    return Py_BuildValue("{sNsNsNsNsd}", "offsets", off, "coarse_peak", cp, "refined_peak", rp,
                         "snr", sn, "latency_ms", latency);
}

PyObject *server_stats(PyObject *, PyObject *args)
{
    PyObject *pyServer;
    if (!PyArg_ParseTuple(args, "O!:server_stats", &PyCapsule_Type, &pyServer)) return nullptr;
    Server *s = get_server(pyServer);
    if (!s) return nullptr;
    uint64_t pairs, batches, graphs, lanes, tunes;
    if (int rc = ampcor_hip_server_stats(s->server, &pairs, &batches, &graphs, &lanes, &tunes)) return raise(rc);
    return Py_BuildValue("{sKsKsKsKsK}", "pairs", (unsigned long long)pairs, "batches",
                         (unsigned long long)batches, "graph_launches", (unsigned long long)graphs,
                         "lanes", (unsigned long long)lanes, "tunes", (unsigned long long)tunes);
}

PyObject *autotune(PyObject *, PyObject *args)
{
    PyObject *pyWorker;
    int reps = 3;
    if (!PyArg_ParseTuple(args, "O!|i:autotune", &PyCapsule_Type, &pyWorker, &reps)) return nullptr;
    Worker *w = get_worker(pyWorker);
    if (!w) return nullptr;
    int rc;
    w->busy = true;
    Py_BEGIN_ALLOW_THREADS
    rc = ampcor_hip_plan_autotune(w->plan, reps);
    Py_END_ALLOW_THREADS
    w->busy = false;
    if (rc) return raise(rc);
    Py_RETURN_NONE;
}

PyObject *device_count(PyObject *, PyObject *)
{
    int n = 0;
    if (int rc = ampcor_hip_device_count(&n)) return raise(rc);
    return PyLong_FromLong(n);
}

PyMethodDef methods[] = {
    {"sequential", (PyCFunction)(void (*)(void))sequential, METH_VARARGS | METH_KEYWORDS,
     "instantiate a new sequential correlation worker"},
    {"addReference", addReference, METH_VARARGS, "process a reference tile"},
    {"addTarget", addTarget, METH_VARARGS, "process a target tile"},
    {"addPairs", addPairs, METH_VARARGS, "register a run of pairs from int64 corner arrays"},
    {"adjust", (PyCFunction)(void (*)(void))adjust, METH_VARARGS | METH_KEYWORDS,
     "perform pixel adjustments to the registration map"},
    {"products", products, METH_VARARGS, "offsets, correlation peaks and SNR of the last adjust"},
    {"dims", dims, METH_VARARGS, "(coarse, refined, zoomed) surface edges"},
    {"stage_gamma", stage_gamma, METH_VARARGS, "coarse correlation surfaces of the last adjust"},
    {"stage_gamma_refined", stage_gamma_refined, METH_VARARGS, "refined correlation surfaces"},
    {"stage_locations", stage_locations, METH_VARARGS, "coarse, window and zoomed peak indices"},
    {"fp8_guard", fp8_guard, METH_VARARGS, "FP8 peak guard counts of the last adjust"},
    {"stage_zoom", stage_zoom, METH_VARARGS, "run the zoom stage on given refined surfaces"},
    {"stage_upsample", (PyCFunction)(void (*)(void))stage_upsample, METH_VARARGS | METH_KEYWORDS,
     "run the target upsampling on given tiles"},
    {"stage_contract", stage_contract, METH_VARARGS, "run one contraction on given operands"},
    {"stage_window_sums", stage_window_sums, METH_VARARGS, "FP64 window sums of given windows"},
    {"stage_amplitudes", stage_amplitudes, METH_VARARGS, "the last FP32 adjust's amplitudes"},
    {"zoom_stats", zoom_stats, METH_VARARGS, "the zoom screen's exact re-evaluations"},
    {"set_timing", set_timing, METH_VARARGS, "enable per-stage device timing"},
    {"stage_times", stage_times, METH_VARARGS, "per-stage device milliseconds of the last adjust"},
    {"raster", (PyCFunction)(void (*)(void))raster, METH_VARARGS | METH_KEYWORDS,
     "upload a raster from a buffer"},
    {"slc_map", (PyCFunction)(void (*)(void))slc_map, METH_VARARGS | METH_KEYWORDS,
     "map an SLC raster file of a given shape onto the device"},
    {"slc_pixelSize", slc_pixelSize, METH_NOARGS, "the size of an SLC pixel"},
    {"slc_getitem", slc_getitem, METH_VARARGS, "fetch the pixel at the given offset"},
    {"refine", refine, METH_VARARGS,
     "sub-pixel refinement of the registration map (adjust already refines)"},
    {"device_count", device_count, METH_NOARGS, "number of visible HIP devices"},
    {"autotune", autotune, METH_VARARGS, "time candidate tilings on the last adjust's data"},
    {"server", (PyCFunction)(void (*)(void))server, METH_VARARGS | METH_KEYWORDS,
     "start a continuous-batching server"},
    {"submit", (PyCFunction)(void (*)(void))submit, METH_VARARGS | METH_KEYWORDS,
     "submit an offset-tracking job; returns its id"},
    {"wait", (PyCFunction)(void (*)(void))wait, METH_VARARGS | METH_KEYWORDS,
     "wait for a job; returns its products and latency"},
    {"server_stats", server_stats, METH_VARARGS, "pairs, batches, graph launches, lanes, tunes"},
    {nullptr, nullptr, 0, nullptr},
};

PyModuleDef module_def = {
    PyModuleDef_HEAD_INIT, "ampcor_hip",
    "ampcor_hip: the MI355X batched tile correlator", -1, methods,
};

}  // namespace

PyMODINIT_FUNC PyInit_ampcor_hip(void)
{
    PyObject *m = PyModule_Create(&module_def);
    if (!m) return nullptr;
    g_error = PyErr_NewException("ampcor_hip.Error", nullptr, nullptr);
    Py_XINCREF(g_error);
    if (PyModule_AddObject(m, "Error", g_error) < 0) {
        Py_XDECREF(g_error);
        Py_DECREF(m);
        return nullptr;
    }
    return m;
}
