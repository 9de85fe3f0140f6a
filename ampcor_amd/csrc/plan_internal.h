// plan_internal.h -- plan and raster internals shared by plan.hip and server.hip.
#pragma once

#include "ampcor_hip.h"
#include "kernels.h"

#include <string>
#include <vector>

struct ampcor_hip_raster {
    int device;
    int dtype;
    const void *data;
    size_t lines, samples;
    bool owned;
};

namespace ampcor_hip {

// records a failure for ampcor_hip_last_error() and returns the code
int fail(int code, const std::string &message);
int hip_fail(hipError_t e, const char *what);

#define HIP_TRY(call, what)                                  \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess) return hip_fail(_e, (what));   \
    } while (0)

// where a plan's device buffers come from: hipMalloc (plans), or carved out of
// a fixed region (the allocation-free entry point)
struct Carver {
    virtual ~Carver() = default;
    virtual hipError_t get(void **ptr, size_t bytes) = 0;
};

struct MallocCarver : Carver {
    hipError_t get(void **ptr, size_t bytes) override { return hipMalloc(ptr, bytes); }
};

// consecutive 256-byte-aligned pieces of [base, base + cap); a null base only
// counts (used = the bytes a plan needs)
struct ScratchCarver : Carver {
    char *base;
    size_t cap, used = 0;
    ScratchCarver(void *b, size_t c) : base(static_cast<char *>(b)), cap(c) {}
    hipError_t get(void **ptr, size_t bytes) override
    {
        used = (used + 255) & ~(size_t)255;
        if (bytes > cap || used > cap - bytes) return hipErrorOutOfMemory;
        *ptr = base ? base + used : nullptr;
        used += bytes;
        return hipSuccess;
    }
};

template <typename T>
struct DeviceBuffer {
    T *ptr = nullptr;
    size_t count = 0;
    hipError_t alloc(size_t n)
    {
        count = n;
        return hipMalloc((void **)&ptr, n * sizeof(T) > 0 ? n * sizeof(T) : 1);
    }
    hipError_t carve(Carver &c, size_t n)
    {
        count = n;
        return c.get((void **)&ptr, n * sizeof(T) > 0 ? n * sizeof(T) : 1);
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
    }
};

constexpr int kStages = 16;  // room for every stage of any form
constexpr size_t kMaxPlanPairs = (size_t)1 << 21;  // pairs one plan may hold

// makes the plan's device current for the duration of a call
struct DeviceGuard {
    int previous = -1;
    explicit DeviceGuard(int device)
    {
        if (hipGetDevice(&previous) != hipSuccess) previous = -1;
        if (previous != device) (void)hipSetDevice(device);
    }
    ~DeviceGuard()
    {
        if (previous >= 0) (void)hipSetDevice(previous);
    }
};

}  // namespace ampcor_hip

struct ampcor_hip_plan {
    using PairDesc = ampcor_hip::PairDesc;
    int device = 0;
    int precision = AMPCOR_HIP_FP32;
    int dtype = -1;
    size_t pairs = 0;
    int C = 0, W = 0, m = 0, P = 0, rf = 0, rm = 0, E = 0, Cr = 0, Wr = 0, R = 0, zf = 0, Z = 0;
    hipStream_t stream = nullptr;

    std::vector<PairDesc> desc;
    std::vector<unsigned char> have_ref, have_tgt;
    bool desc_dirty = true;
    bool adjusted = false;

    ampcor_hip::CorrelateShape coarse_shape{}, refined_shape{};
    ampcor_hip::MixedShape mixed_coarse{}, mixed_refined{};
    ampcor_hip::Fp8Shape fp8_coarse{};
    bool fp8 = false;    // FP8 coarse contraction (refined stage as FP16)
    bool mixed = false;
    bool split = false;  // hi/lo operands in the refined contraction
    bool fp16 = false;

    ampcor_hip::DeviceBuffer<PairDesc> d_desc;
    ampcor_hip::DeviceBuffer<float> amp_ref, amp_tgt, rT, partial, gamma;
    ampcor_hip::DeviceBuffer<double2> sums, sums_r;
    ampcor_hip::DeviceBuffer<double2> stats, stats_r;
    ampcor_hip::DeviceBuffer<int2> loc, window, zoom_loc;
    ampcor_hip::DeviceBuffer<float> peak, snr, peak_r, offsets;
    ampcor_hip::DeviceBuffer<float> amp_ref_r, amp_tgt_r, rT_r, partial_r, gamma_r;
    ampcor_hip::DeviceBuffer<float> U_ref, U_tgt, g_ref, g_tgt;
    ampcor_hip::DeviceBuffer<double> D;
    // the zoom's FP16 screen: Dirichlet fragments, its error bound
    // (ecoef max|g| + eabs) and counts of the cells it sent to exact evaluation
    ampcor_hip::DeviceBuffer<uint16_t> zfrag, zfragp;
    ampcor_hip::ZoomBounds zoom_bounds{};
    ampcor_hip::DeviceBuffer<unsigned> zstats;
    ampcor_hip::DeviceBuffer<double> zthr;  // the zoom kernels' hand-off, per pair
    ampcor_hip::DeviceBuffer<int32_t> zwin;
    ampcor_hip::DeviceBuffer<int> zflag;
    ampcor_hip::DeviceBuffer<int> zncand;
    ampcor_hip::DeviceBuffer<int> zcand;
    ampcor_hip::ZoomScratch zoom_scratch() const
    {
        return {zthr.ptr, zwin.ptr, zflag.ptr, zncand.ptr, zcand.ptr};
    }
    ampcor_hip::DeviceBuffer<uint16_t> chip_q, tgt_q, chip_q_r, tgt_q_r, chip_lo_r, tgt_lo_r;
    ampcor_hip::DeviceBuffer<uint16_t> hfrag_ref, hfrag_tgt;  // FP16 circulant fragments
    // scratch of the banded upsampler (tiles above 96 samples)
    ampcor_hip::DeviceBuffer<float2> up_T;
    ampcor_hip::DeviceBuffer<double2> up_band, up_rows;
    ampcor_hip::UpsampleScratch up_scratch{};
    ampcor_hip::DeviceBuffer<ampcor_hip::MixedStats> mstats, mstats_r;
    // FP8 peak guard counters: re-scored, moved, over the candidate limit
    ampcor_hip::DeviceBuffer<unsigned> guard;
    ampcor_hip::DeviceBuffer<int> near;  // cells within the guard's margin of each FP8 peak
    // pinned host staging of the offsets for adjust's host copy (allocated on
    // its first use): a copy from pageable memory goes through the runtime's
    // own staging, with a synchronization per chunk
    float *offsets_host = nullptr;

    bool timing = false;
    hipEvent_t ev[ampcor_hip::kStages + 1] = {};
    float stage_ms[ampcor_hip::kStages] = {};
    const char *stage_name[ampcor_hip::kStages] = {};
    int stages_run = 0;
    int last_n = 0;  // pairs of the last enqueue

    void release()
    {
        ampcor_hip::DeviceBuffer<float> *fb[] = {&amp_ref, &amp_tgt, &rT, &partial, &gamma, &peak, &snr,
                                     &peak_r, &offsets, &amp_ref_r, &amp_tgt_r, &rT_r,
                                     &partial_r, &gamma_r, &U_ref, &U_tgt, &g_ref, &g_tgt};
        for (auto *b : fb) b->release();
        if (offsets_host) (void)hipHostFree(offsets_host);
        offsets_host = nullptr;
        sums.release();
        sums_r.release();
        chip_q.release();
        tgt_q.release();
        chip_q_r.release();
        tgt_q_r.release();
        chip_lo_r.release();
        tgt_lo_r.release();
        hfrag_ref.release();
        hfrag_tgt.release();
        up_T.release();
        up_band.release();
        up_rows.release();
        mstats.release();
        mstats_r.release();
        guard.release();
        near.release();
        D.release();
        zfrag.release();
        zfragp.release();
        zstats.release();
        zthr.release();
        zwin.release();
        zflag.release();
        zncand.release();
        zcand.release();
        stats.release();
        stats_r.release();
        loc.release();
        window.release();
        zoom_loc.release();
        d_desc.release();
        for (auto &e : ev)
            if (e) {
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};


namespace ampcor_hip {
// enqueue every stage for the first n registered pairs on the plan's stream;
// uploads the descriptors first when they changed; no synchronization
int plan_enqueue(ampcor_hip_plan *plan, int n);
// validates a geometry and fills a plan's dimensions and tilings (no device
// memory); the arguments are those of ampcor_hip_plan_create
int plan_configure(ampcor_hip_plan *p, int device, size_t pairs, size_t ref_lines,
                   size_t ref_samples, size_t tgt_lines, size_t tgt_samples, size_t refine_margin,
                   size_t refine_factor, size_t zoom_factor, int precision);
// takes every device buffer of a configured plan from `cv` and uploads its
// tables on stream s (synchronizing it); with upload false only carves (a
// ScratchCarver with a null base then counts the bytes)
int plan_build(ampcor_hip_plan *p, Carver &cv, hipStream_t s, bool upload);
// device bytes a plan of this geometry allocates per pair (for admission)
size_t plan_bytes_per_pair(int C, int W, int rf, int rm, int precision);
}  // namespace ampcor_hip
