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

// per-pair constants of the mixed normalizer, unscaled units
struct MixedStats {
    double sumr;     // sum of the quantized chip values (unscaled)
    double sumr2;    // sum of their squares
    double scale_r;  // power-of-two scale of the quantized chip
    double scale_t;  // power-of-two scale of the quantized window
    double mean_t;   // constant removed from the window before rounding
};
// quantization modes of the upsampler: detected FP32 only, or BF16 / FP16
// operands of the mixed contraction as well
enum UpsampleOutput { kUpFP32 = 0, kUpBF16 = 1, kUpFP16 = 2 };
// OCP e4m3 operands (one byte each) of the FP8 coarse contraction; written by
// the mixed gather only
constexpr int kUpFP8 = 3;

// tiling of one correlation launch: placements [u0, u0+ub) x [v0, v0+vb) of one
// pair, chip columns split into `nsplit` ranges, each its own workgroup
struct CorrelateShape {
    bool exact;  // one range; FP32 chains per chip column summed in FP64, normalized in place
    int bu;      // placement rows per lane (register block)
    int ub, vb;  // placement tile extent
    int tiles_u, tiles_v;
    int nsplit;  // chip-column ranges (partial numerators summed by the normalizer)
    int jblock;  // chip columns staged per LDS pass
    int threads;
    size_t lds_bytes;
};

// chooses the tiling for chip C over P x P placements; exact = keep each
// placement's whole contraction in one lane (the FP32 drop-in order);
// bu / nsplit / jblock > 0 request a register block / chip-column split /
// chip columns per LDS pass (tuning)
CorrelateShape choose_correlate_shape(int C, int P, bool exact, int bu = 0, int nsplit = 0,
                                      int jblock = 0);

hipError_t launch_gather_detect(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int C,
                                int W, float *ref_amp, float *tgt_amp);
// exact: the FP32 twin's order (the coarse stage); otherwise any order
hipError_t launch_chip_stats(hipStream_t s, const float *amp, int pairs, int C, float *rT,
                             double2 *stats, bool exact = true);
// FP64 sums of t and t^2 over every C x C placement of each W x W window: [pairs][P][P]
hipError_t launch_window_sums(hipStream_t s, const float *amp, int pairs, int W, int C,
                              double2 *sums);
// exact shapes write gamma (from the window sums and chip statistics, which
// must be given); split shapes write the FP32 partial numerators
hipError_t launch_correlate(hipStream_t s, const CorrelateShape &shape, const float *rT,
                            const float *amp, int pairs, int C, int W, float *partial,
                            const double2 *sums = nullptr, const double2 *stats = nullptr,
                            float *gamma = nullptr);
hipError_t launch_normalize(hipStream_t s, const float *partial, int nsplit, const double2 *sums,
                            const double2 *stats, int pairs, int C, int W, float *gamma);
// window sums and normalization in one pass (W <= 512, P <= 144; tolerance-checked
// stages).  form kFuseFP32: FP32 amplitudes and `stats` (sum r', |r'|);
// kFuseMixed: FP32 amplitudes and `mstats`; kFuseBF16 / kFuseFP16 / kFuseFP8:
// the sums come from the quantized target (the contraction's operand, two
// bytes or one byte per value) and `mstats`
enum { kFuseFP32 = 0, kFuseMixed = 1, kFuseBF16 = 2, kFuseFP16 = 3, kFuseFP8 = 4 };
bool normalize_fuses(int W, int C);
hipError_t launch_normalize_fused(hipStream_t s, int form, const void *target, int pairs, int W,
                                  int C, const float *partial, int nsplit, const double2 *stats,
                                  const MixedStats *mstats, float *gamma);
// near (optional): per pair, the number of cells within tau of the peak
hipError_t launch_coarse_peak(hipStream_t s, const float *gamma, int pairs, int P, int W, int E,
                              int refine_margin, int exclude, int2 *loc, int2 *window,
                              float *peak, float *snr, float tau = 0.0f, int *near = nullptr);
// FP8 coarse peak guard: for the pairs whose peak kernel counted more than
// one cell within peak_guard_margin(C) of the FP8 peak (near), those cells
// are re-scored by the FP64 correlation of the raster amplitudes, and the best
// one becomes the peak location (loc, window, peak and snr rewritten for that
// pair).  counters (3 unsigned, device):
// pairs re-scored, pairs whose location moved, pairs with more candidates than
// the guard re-scores (their FP8 peak is kept); cleared by the launch.
float peak_guard_margin(int C);
hipError_t launch_peak_guard(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int C,
                             int W, int E, int rm, int exclude, const float *gamma, const int *near,
                             int2 *loc, int2 *window, float *peak, float *snr, unsigned *counters);
// band-limited upsampling of complex tiles straight from the rasters, detected:
// role 0 = reference chips (N = C, origin = chip corner), role 1 = target
// windows (N = E, origin = window corner + nudged (top, left)); U is the
// (rf N) x N upsampling matrix stored padded to multiples of 16, zero-filled;
// g (rf = 2 only) is its half-sample row, U[2a + 1][n] = g((a - n) mod N).
// With a quantized output mode (rf = 2, N <= 512) the kernel also writes the
// mixed-contraction operand: role 0 the chip rows q[pair][M][cs] (16 zero
// columns on the left), role 1 the window q[pair][M][M]; and its share of
// the pair's MixedStats.  amp may be null then (role 1 too, when the FP16
// fragments are given and the window sums come from the quantized target).
bool upsample_fuses_quantization(int N, int rf);
// qlo (optional): second quantized plane holding the rounding remainder of q
// (the split hi/lo operands), same layout as q.
// hfrag (optional, quantized modes without qlo): the half-sample circulant as
// FP16 MFMA fragments (upsample_fragments); the products then run on FP16
// matrix cores instead of FP32 ones
int upsample_fragment_count(int N);
// whether a tile edge N can use them (up to 96 for complex rasters, 160 for real)
bool upsample_uses_fragments(int N);
// tiles with 96 < N <= 512 (rf = 2) upsample in two banded passes through
// global scratch: T [pairs][N][N], band and row statistics [pairs][ceil(N / 16)]
struct UpsampleScratch {
    float2 *T;
    double2 *band, *rows;
};
bool upsample_needs_scratch(int N, int rf);
hipError_t launch_upsample(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int role,
                           const int2 *window, int N, int rf, const float *U, const float *g,
                           float *amp, int qmode = kUpFP32, uint16_t *q = nullptr, int cs = 0,
                           MixedStats *stats = nullptr, uint16_t *qlo = nullptr,
                           const uint16_t *hfrag = nullptr,
                           const UpsampleScratch *scratch = nullptr);
// mixed coarse stage in one pass when the tiles fit in LDS: gather, detect,
// write the FP32 window (for the window sums) and both quantized operands
bool gather_prep_fits(int C, int W);
// qmode kUpBF16 / kUpFP16 / kUpFP8; with kUpFP8 chip_q and tgt_q hold bytes
// (rows of CS bytes and W x W bytes)
hipError_t launch_gather_prep_mixed(hipStream_t s, const PairDesc *desc, int pairs, int dtype,
                                    int C, int W, int CS, int qmode, float *tgt_amp,
                                    void *chip_q, void *tgt_q, MixedStats *stats);
// ---- mixed precision (mixed.hip) ----

// tiling of the MFMA contraction: 16 x 16 placement tiles, nk K-steps of 32
// target columns per tile, workgroups of wg_ut waves x wg_vt v-tiles
struct MixedShape {
    int nk, cs;
    bool split;
    int wg_ut, wg_vt, tiles_u, tiles_v, nsplit;
    size_t lds_bytes;
    bool stream;   // the streaming form: all chip rows per workgroup, staged chunk by chunk
    // the packed form (P = 32 NB + 1): chip rows per chunk, 0 otherwise; one
    // workgroup per pair of wg_ut = NB x tiles_v waves, each wg_vt v-tiles wide
    int packed;
    int wpe;   // packed form: waves per SIMD its registers are held to (2 or 3)
};
// nsplit > 0 asks for at least that many chip-row ranges (tuning)
// split: hi/lo operands (three products per pair of elements), twice the LDS
// max_vt > 0 caps the v-tiles per workgroup (tuning)
MixedShape choose_mixed_shape(int C, int P, int nsplit = 0, bool split = false, int max_vt = 0);
// the streaming form's shape (plain operands only); nk = 0 where it does not apply
MixedShape choose_mixed_stream_shape(int C, int P, int max_vt = 0);
// the packed form's shape (P = 32 NB + 1, chips of whole 16-column blocks;
// split: hi / lo operands, three passes; nsplit: chip-row ranges, a power of
// two whose ranges hold whole chunks); nk = 0 where it does not apply
MixedShape choose_mixed_packed_shape(int C, int P, int max_vt = 0, int wpe = 0,
                                     bool split = false, int nsplit = 1);
// a packed shape launched over `pairs` pairs: its chip-row ranges raised (up to
// maxsplit, whole chunks each) until the launch has enough waves to fill the
// chip; other shapes unchanged
MixedShape packed_fill_ranges(const MixedShape &sh, int C, int pairs, int maxsplit);
// chip_lo / tgt_lo (optional): the rounding remainders, for the split form
hipError_t launch_prep_mixed(hipStream_t s, const float *ref_amp, const float *tgt_amp, int pairs,
                             int C, int W, int CS, bool fp16, uint16_t *chip_q, uint16_t *tgt_q,
                             MixedStats *stats, uint16_t *chip_lo = nullptr,
                             uint16_t *tgt_lo = nullptr);
hipError_t launch_correlate_mixed(hipStream_t s, const MixedShape &shape, const uint16_t *chip_q,
                                  const uint16_t *tgt_q, int pairs, int C, int W, bool fp16,
                                  float *partial, const uint16_t *chip_lo = nullptr,
                                  const uint16_t *tgt_lo = nullptr);
hipError_t launch_normalize_mixed(hipStream_t s, const float *partial, int nsplit,
                                  const double2 *sums, const MixedStats *stats, int pairs, int C,
                                  int W, float *gamma);

// ---- FP8 coarse contraction (fp8.hip) ----

// tiling of the FP8 contraction (v_mfma_f32_16x16x128_f8f6f4): K-steps of 128
// = eight chip rows x 16 target columns; chip operand rows of cs bytes (16
// zero bytes, C values, zeros); chip rows split into nsplit ranges of ilen (a
// multiple of 8, zero past C)
struct Fp8Shape {
    int nk, cs;
    int wg_ut, wg_vt, tiles_u, tiles_v, nsplit, ilen;
    size_t lds_bytes;
};
Fp8Shape choose_fp8_shape(int C, int P, int nsplit = 0, int max_vt = 0);
hipError_t launch_correlate_fp8(hipStream_t s, const Fp8Shape &shape, const uint8_t *chip_q,
                                const uint8_t *tgt_q, int pairs, int C, int W, float *partial);

// whether the zoom of an R x R refined surface by zf fits one workgroup's LDS
bool zoom_fits(int R, int zf);
// the zoom's FP16 Dirichlet fragments (halves): for row tile zt < ceil(zf R / 16),
// K-chunk kc < ceil(R / 32), lane l, value e: D(16 zt + l % 16 - zf b) with
// b = 32 kc + 8 (l / 16) + e (dfrag), or, for the transposed screen's A
// operand, b = 32 kc + 4 (l / 16) + e for e < 4 and 32 kc + 16 + 4 (l / 16) +
// e - 4 otherwise (pfrag); zero past zf R rows or R columns
int zoom_fragment_count(int R, int zf);
// cells outside its window that the zoom's screen lists per pair for exact
// evaluation (a pair with more gets the full search)
constexpr int kZoomList = 512;
// per-pair scratch of the zoom's kernels: [pairs] each
struct ZoomScratch {
    double *thr;
    int32_t *win;  // [pairs][4]
    int *flag;
    int *ncand;
    int *cand;     // [pairs][kZoomList]
};
// the caller's bounds on the zoom screen's FP16 error (g = the refined surface):
// |f_screen - f| <= ecoef max|g| + eabs over the two products, |H_screen - H|
// <= hcoef max|g| + habs after the first; lambda = max_p sum_a |D(p - zf a)|
struct ZoomBounds {
    double ecoef, eabs;
    float lambda, hcoef, habs;
};
// zoom over the whole zf R x zf R grid: the exact FP64 maximum (first in
// row-major order), certified by an FP16 screen; stats (optional, 3 unsigned):
// cells evaluated exactly outside the windows, pairs that had any, pairs given
// the full search
hipError_t launch_zoom(hipStream_t s, const float *gamma_refined, int pairs, int R, int zf,
                       const double *dirichlet, const uint16_t *dfrag, const uint16_t *pfrag,
                       const ZoomBounds &bounds, const int2 *window, const float *coarse_peak,
                       int margin, int rf, float *offsets, float *refined_peak, int2 *zoom_loc,
                       const ZoomScratch &scratch, unsigned *stats = nullptr);

// pairs that one launch may take when each pair brings `items` work-items: a
// dispatch's work-items (workgroups x threads) must fit in 32 bits
inline int launch_pair_cap(size_t items)
{
    const size_t cap = ((size_t)0xffffffffu - 1023) / (items ? items : 1);
    return cap < 1 ? 1 : cap > (size_t)0x7fffffff ? 0x7fffffff : (int)cap;
}

// the logical index of workgroup b of n: the hardware hands consecutive
// workgroups to the eight XCDs in turn, and this gives each XCD a contiguous
// run of logical indices, so the workgroups that read one pair's operands (or
// neighbouring pairs of the scene) meet in the same L2.  It is a bijection on
// [0, n), so results do not depend on where the workgroups actually run.
__device__ inline int xcd_block(int b, int n)
{
    const int x = b & 7, q = n >> 3, r = n & 7;
    return x * q + min(x, r) + (b >> 3);
}

}  // namespace ampcor_hip
