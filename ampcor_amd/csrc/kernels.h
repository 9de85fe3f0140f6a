// kernels.h -- device kernels of the gfx950 tile correlator and their launchers.
//
// Every launcher enqueues on the given stream and returns the launch status;
// none synchronizes.  Buffer layouts are pair-major and row-major throughout.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ampcor_hip {

// one tile pair: where its reference chip and target window start in their rasters
struct PairDesc {
    const void *ref;     // reference raster base (device)
    const void *tgt;     // target raster base (device)
    int32_t ref_stride;  // samples per raster line
    int32_t tgt_stride;
    int32_t ref_line, ref_sample;  // upper-left corner of the C x C chip
    int32_t tgt_line, tgt_sample;  // upper-left corner of the W x W window
};

// tiling of one correlation launch: placements [u0, u0+ub) x [v0, v0+vb) of one
// pair, chip columns split into `nsplit` ranges, each its own workgroup
struct CorrelateShape {
    int bu;      // placement rows per lane (register block)
    int ub, vb;  // placement tile extent
    int tiles_u, tiles_v;
    int nsplit;  // chip-column ranges (partial numerators summed by the normalizer)
    int threads;
    size_t lds_bytes;
};

// chooses the tiling for chip C over P x P placements; exact = keep each
// placement's whole contraction in one lane (the FP32 drop-in order);
// bu / nsplit > 0 request a register block / chip-column split (tuning)
CorrelateShape choose_correlate_shape(int C, int P, bool exact, int bu = 0, int nsplit = 0);

hipError_t launch_gather_detect(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int C,
                                int W, float *ref_amp, float *tgt_amp);
hipError_t launch_chip_stats(hipStream_t s, const float *amp, int pairs, int C, float *rT,
                             double2 *stats);
// FP64 sums of t and t^2 over every C x C placement of each W x W window: [pairs][P][P]
hipError_t launch_window_sums(hipStream_t s, const float *amp, int pairs, int W, int C,
                              double2 *sums);
hipError_t launch_correlate(hipStream_t s, const CorrelateShape &shape, const float *rT,
                            const float *amp, int pairs, int C, int W, float *partial);
hipError_t launch_normalize(hipStream_t s, const float *partial, int nsplit, const double2 *sums,
                            const double2 *stats, int pairs, int C, int W, float *gamma);
hipError_t launch_coarse_peak(hipStream_t s, const float *gamma, int pairs, int P, int W, int E,
                              int refine_margin, int exclude, int2 *loc, int2 *window,
                              float *peak, float *snr);
// band-limited upsampling of complex tiles straight from the rasters, detected:
// role 0 = reference chips (N = C, origin = chip corner), role 1 = target
// windows (N = E, origin = window corner + nudged (top, left)); U is the
// (rf N) x N upsampling matrix stored padded to multiples of 16, zero-filled;
// g (rf = 2 only) is its half-sample row, U[2a + 1][n] = g((a - n) mod N)
hipError_t launch_upsample(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int role,
                           const int2 *window, int N, int rf, const float *U, const float *g,
                           float *amp);
// ---- mixed precision (mixed.hip) ----
// per-pair constants of the mixed normalizer, unscaled units
struct MixedStats {
    double sumr;    // sum of the quantized zero-mean chip
    double normr;   // its L2 norm
    double inv_ab;  // 1 / (chip scale * target scale)
    double mean_t;  // window mean removed from the quantized target
};
// tiling of the MFMA contraction: 16 x 16 placement tiles, nk K-steps of 16
// target columns per tile, workgroups of wg_ut waves x wg_vt v-tiles
struct MixedShape {
    int nk, cs;
    int wg_ut, wg_vt, tiles_u, tiles_v, nsplit;
    size_t lds_bytes;
};
MixedShape choose_mixed_shape(int C, int P);
hipError_t launch_prep_mixed(hipStream_t s, const float *ref_amp, const float *tgt_amp, int pairs,
                             int C, int W, int CS, bool fp16, uint16_t *chip_q, uint16_t *tgt_q,
                             MixedStats *stats);
hipError_t launch_correlate_mixed(hipStream_t s, const MixedShape &shape, const uint16_t *chip_q,
                                  const uint16_t *tgt_q, int pairs, int C, int W, bool fp16,
                                  float *partial);
hipError_t launch_normalize_mixed(hipStream_t s, const float *partial, int nsplit,
                                  const double2 *sums, const MixedStats *stats, int pairs, int C,
                                  int W, float *gamma);

hipError_t launch_zoom(hipStream_t s, const float *gamma_refined, int pairs, int R, int zf,
                       const double *dirichlet, const int2 *window, const float *coarse_peak,
                       int margin, int rf, float *offsets, float *refined_peak, int2 *zoom_loc);

}  // namespace ampcor_hip
