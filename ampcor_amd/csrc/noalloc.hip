// noalloc.hip -- ampcor_hip_adjust_noalloc: a whole adjust (plan, pairs,
// correlation, refinement) that allocates no device memory.
//
// Every intermediate buffer and table of the pipeline is carved out of one
// file-scope device array, `scratch` (AMPCOR_HIP_NOALLOC_SCRATCH_BYTES per
// device, held by every device the library runs on until the process ends).
// In host code the array's name is not a device address: its address on each
// device comes from hipGetSymbolAddress, called once per device with that
// device current, and only that address is used.  Calls on one device share
// its array and overlap at their peril (they race); each call writes every
// buffer before its kernels read it and the kernels clear their own counters,
// so nothing depends on what an earlier call left behind.
#include "plan_internal.h"

#include <mutex>
#include <string>
#include <vector>

using namespace ampcor_hip;

namespace {

constexpr size_t kScratchBytes = AMPCOR_HIP_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
std::vector<void *> g_scratch;  // per device: the array's address there

// the scratch array's address on `device` (current), looked up on first use
int scratch_on(int device, void **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return hip_fail(e, "locating the scratch array");
        g_scratch[device] = ptr;
    }
    *out = g_scratch[device];
    return AMPCOR_HIP_OK;
}

// device bytes a configured plan carves (tables included)
int plan_scratch_bytes(ampcor_hip_plan *p, size_t *bytes)
{
    ScratchCarver count(nullptr, (size_t)-1);
    if (int rc = plan_build(p, count, nullptr, false)) return rc;
    *bytes = count.used;
    return AMPCOR_HIP_OK;
}

}  // namespace

extern "C" int ampcor_hip_adjust_noalloc(
    int device, void *stream, int dtype, const void *reference, size_t ref_lines,
    size_t ref_samples, const void *target, size_t tgt_lines, size_t tgt_samples, size_t pairs,
    const int64_t *ref_begin, const int64_t *tgt_begin, size_t chip, size_t window,
    size_t refine_margin, size_t refine_factor, size_t zoom_factor, int precision, float *offsets,
    float *coarse_peak, float *refined_peak, float *snr)
{
    if (!reference || !target || !offsets || !coarse_peak || !refined_peak || !snr)
        return fail(AMPCOR_HIP_EINVAL, "null raster or output pointer");
    if (dtype != AMPCOR_HIP_COMPLEX64 && dtype != AMPCOR_HIP_FLOAT32)
        return fail(AMPCOR_HIP_EINVAL, "unknown raster pixel type");
    // the plan lives on the host only: its device buffers are carved from scratch
    ampcor_hip_plan p;
    if (int rc = plan_configure(&p, device, pairs, chip, chip, window, window, refine_margin,
                                refine_factor, zoom_factor, precision))
        return rc;
    size_t need = 0;
    if (int rc = plan_scratch_bytes(&p, &need)) return rc;
    if (need > kScratchBytes) {
        // the largest batch of this geometry that fits, from the bytes of one and two pairs
        ampcor_hip_plan one, two;
        size_t b1 = 0, b2 = 0;
        std::string most = "?";
        if (!plan_configure(&one, device, 1, chip, chip, window, window, refine_margin,
                            refine_factor, zoom_factor, precision) &&
            !plan_configure(&two, device, 2, chip, chip, window, window, refine_margin,
                            refine_factor, zoom_factor, precision) &&
            !plan_scratch_bytes(&one, &b1) && !plan_scratch_bytes(&two, &b2) && b2 > b1 &&
            b1 <= kScratchBytes)
            most = std::to_string(1 + (kScratchBytes - b1) / (b2 - b1));
        return fail(AMPCOR_HIP_EINVAL,
                    std::to_string(pairs) + " pairs of this geometry need " + std::to_string(need) +
                        " bytes of device scratch; the allocation-free entry point has " +
                        std::to_string(kScratchBytes) + " (at most " + most +
                        " pairs a call): split the batch");
    }
    DeviceGuard guard(device);
    void *base = nullptr;
    if (int rc = scratch_on(device, &base)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    p.stream = s;
    ScratchCarver cv(base, kScratchBytes);
    if (int rc = plan_build(&p, cv, s, true)) return rc;
    // the caller's outputs take the place of the plan's
    p.offsets.ptr = offsets;
    p.peak.ptr = coarse_peak;
    p.peak_r.ptr = refined_peak;
    p.snr.ptr = snr;
    // the pairs, checked against the rasters as ampcor_hip_add_pairs checks them
    ampcor_hip_raster ref{device, dtype, reference, ref_lines, ref_samples, false};
    ampcor_hip_raster tgt{device, dtype, target, tgt_lines, tgt_samples, false};
    if (int rc = ampcor_hip_add_pairs(&p, &ref, &tgt, 0, pairs, ref_begin, tgt_begin)) return rc;
    if (int rc = plan_enqueue(&p, (int)pairs)) return rc;
    // the descriptors are copied from the plan's host vector: wait before it goes
    HIP_TRY(hipStreamSynchronize(s), "adjust");
    return AMPCOR_HIP_OK;
}
