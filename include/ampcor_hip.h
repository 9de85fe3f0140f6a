/*
 * ampcor_hip.h -- C ABI of the MI355X (gfx950) batched tile correlator.
 *
 * This is the boundary a HIP correlator plugs into beside ampcor's Sequential
 * and CUDA workers.  Every entry point returns an int status (0 = success) and
 * never throws; ampcor_hip_last_error() describes the most recent failure on
 * the calling thread.  No torch or C++ types cross the boundary.
 *
 * Mapping to the reference interface (paths relative to the ampcor tree):
 *
 *   ampcor_hip_plan_create   <- ampcor_cuda.sequential(pairs, (refLines, refSamples),
 *                               (tgtLines, tgtSamples), refineMargin, refineFactor,
 *                               zoomFactor): ext/ampcor_cuda/sequential.cc:40-73,
 *                               worker ctor lib/libampcor_cuda/correlators/Sequential.icc:173-234
 *   ampcor_hip_add_reference <- ampcor_cuda.addReference(worker, slc, idx, begin, end):
 *                               ext/ampcor_cuda/sequential.cc:85-143,
 *                               Sequential::addReferenceTile Sequential.icc:15-28
 *   ampcor_hip_add_target    <- ampcor_cuda.addTarget(...): ext/ampcor_cuda/sequential.cc:155-213,
 *                               Sequential::addTargetTile Sequential.icc:31-44
 *   ampcor_hip_add_pairs     <- the per-pair loop of pkg/ampcor/correlators/CUDA.py:50-54,
 *                               batched into one call
 *   ampcor_hip_adjust        <- ampcor_cuda.adjust(worker): ext/ampcor_cuda/sequential.cc:225-271,
 *                               Sequential::adjust Sequential.icc:47-141
 *   ampcor_hip_plan_destroy  <- the worker capsule destructor ext/ampcor_cuda/sequential.cc:274-286
 *   ampcor_hip_raster_*      <- the SLC raster the tiles are sliced from:
 *                               lib/libampcor/dom/SLC.h:18-80, ampcor.slc_map ext/ampcor/slc.cc:50-93;
 *                               here the raster is uploaded once and stays in device memory
 *   ampcor_hip_products,
 *   ampcor_hip_stage_*       no counterpart: the reference returns only the offsets and
 *                            exposes its stages as C++ methods to its tests
 *                            (lib/libampcor_cuda/correlators/Sequential.h:69-113)
 */
#ifndef AMPCOR_HIP_H
#define AMPCOR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define AMPCOR_HIP_OK 0
#define AMPCOR_HIP_EINVAL 1  /* bad argument (shape, index, slice, null pointer) */
#define AMPCOR_HIP_ENOMEM 2  /* device or host allocation failed */
#define AMPCOR_HIP_EDEVICE 3 /* HIP runtime error, or no device */
#define AMPCOR_HIP_ESTATE 4  /* call out of order (e.g. products before adjust) */

/* arithmetic of the correlation contraction */
#define AMPCOR_HIP_FP32 0 /* plain FP32: the drop-in form */
#define AMPCOR_HIP_BF16 1 /* mixed: BF16 MFMA contraction, FP32 accumulate */
#define AMPCOR_HIP_FP16 2 /* mixed: FP16 MFMA contraction, FP32 accumulate */
/* mixed, with split hi/lo operands (three products) in the refined re-correlation */
#define AMPCOR_HIP_BF16X3 3
#define AMPCOR_HIP_FP16X3 4
/* mixed: the coarse contraction on FP8 (OCP e4m3) matrix cores, the refined
 * re-correlation as AMPCOR_HIP_FP16 */
#define AMPCOR_HIP_FP8 5

/* raster pixel types */
#define AMPCOR_HIP_COMPLEX64 0 /* SLC: interleaved float re, im */
#define AMPCOR_HIP_FLOAT32 1   /* real amplitude */

typedef struct ampcor_hip_raster ampcor_hip_raster;
typedef struct ampcor_hip_plan ampcor_hip_plan;

/* description of the most recent failure on this thread ("" if none) */
const char *ampcor_hip_last_error(void);
/* number of visible devices */
int ampcor_hip_device_count(int *count);

/* -- rasters --------------------------------------------------------------- */

/* copy a host raster (lines x samples, row-major) into device memory */
int ampcor_hip_raster_create(int device, int dtype, const void *host, size_t lines,
                             size_t samples, ampcor_hip_raster **out);
/* wrap device memory the caller owns and keeps alive; nothing is copied */
int ampcor_hip_raster_wrap(int device, int dtype, const void *device_ptr, size_t lines,
                           size_t samples, ampcor_hip_raster **out);
int ampcor_hip_raster_shape(const ampcor_hip_raster *raster, size_t *lines, size_t *samples);
/* copy `count` pixels starting at row-major pixel `offset` to host memory (8 bytes
 * a pixel for complex64, 4 for float32); AMPCOR_HIP_EINVAL past the raster's end.
 * Replaces the SLC element read of ext/ampcor/slc.cc:105-140 (slc_getitem) and
 * lib/libampcor/dom/SLC.icc (operator[]). */
int ampcor_hip_raster_read(const ampcor_hip_raster *raster, size_t offset, size_t count,
                           void *host);
int ampcor_hip_raster_destroy(ampcor_hip_raster *raster);

/* -- plans ----------------------------------------------------------------- */

/*
 * A plan for `pairs` tile pairs: reference chips of ref_lines x ref_samples,
 * target windows of tgt_lines x tgt_samples (both square, window - chip even,
 * (window - chip) / 2 >= refine_margin).  Device buffers are allocated here
 * and reused by every adjust.  A plan is not thread-safe.
 */
int ampcor_hip_plan_create(int device, size_t pairs, size_t ref_lines, size_t ref_samples,
                           size_t tgt_lines, size_t tgt_samples, size_t refine_margin,
                           size_t refine_factor, size_t zoom_factor, int precision,
                           ampcor_hip_plan **out);
int ampcor_hip_plan_destroy(ampcor_hip_plan *plan);
/* number of pairs the plan was created for */
int ampcor_hip_plan_pairs(const ampcor_hip_plan *plan, size_t *pairs);

/* register the reference chip of pair idx: raster rows [begin_line, end_line),
 * columns [begin_sample, end_sample); the extent must equal the chip shape */
int ampcor_hip_add_reference(ampcor_hip_plan *plan, const ampcor_hip_raster *raster, size_t idx,
                             size_t begin_line, size_t begin_sample, size_t end_line,
                             size_t end_sample);
/* register the target search window of pair idx */
int ampcor_hip_add_target(ampcor_hip_plan *plan, const ampcor_hip_raster *raster, size_t idx,
                          size_t begin_line, size_t begin_sample, size_t end_line,
                          size_t end_sample);
/* register pairs [first, first + count) at once: ref_begin / tgt_begin hold
 * count (line, sample) upper-left corners */
int ampcor_hip_add_pairs(ampcor_hip_plan *plan, const ampcor_hip_raster *reference,
                         const ampcor_hip_raster *target, size_t first, size_t count,
                         const int64_t *ref_begin, const int64_t *tgt_begin);

/* correlate every registered pair; offsets (host, 2 * pairs floats, pair-major
 * (dLine, dSample)) may be NULL to leave the products on the device */
int ampcor_hip_adjust(ampcor_hip_plan *plan, float *offsets);

/* copy the products of the last adjust to host arrays (any may be NULL):
 * offsets [pairs][2], coarse correlation peak [pairs], zoomed refined peak
 * [pairs], coarse SNR [pairs] */
int ampcor_hip_products(ampcor_hip_plan *plan, float *offsets, float *coarse_peak,
                        float *refined_peak, float *snr);

/* -- stage access, for tests and profiling ---------------------------------- */

/* surface dimensions: coarse P = window - chip + 1, refined R = 2 rf rm + 1 */
int ampcor_hip_plan_dims(const ampcor_hip_plan *plan, int *coarse_dim, int *refined_dim,
                         int *zoomed_dim);
/* the refined window's source and result edges: E = chip + 2 refine margin
 * (the expanded target window), rf E */
int ampcor_hip_plan_tile_dims(const ampcor_hip_plan *plan, size_t *expanded, size_t *refined_window);
/* coarse correlation surfaces [pairs][P][P] */
int ampcor_hip_stage_gamma(ampcor_hip_plan *plan, float *gamma);
/* refined correlation surfaces [pairs][R][R] */
int ampcor_hip_stage_gamma_refined(ampcor_hip_plan *plan, float *gamma);
/* integer locations [pairs][2]: coarse argmax (row, col), nudged window
 * (top, left), zoomed argmax (p, q) */
int ampcor_hip_stage_locations(ampcor_hip_plan *plan, int32_t *coarse, int32_t *window,
                               int32_t *zoomed);
/* run the target upsampling of the plan's adjust alone on host tiles
 * [count][E][E] (E = chip + 2 refine margin; complex64 or float32, count <=
 * pairs): amp receives |y| [count][rf E][rf E] of the band-limited x rf
 * upsampling y of each tile -- for the plain BF16/FP16/FP8 forms through their
 * FP16 matrix-core upsampler (k_upsample2h, up to 96 complex or 160 real
 * samples), otherwise the FP32 one (k_upsample2 up to 96, the banded two-pass
 * form above, the dense form for rf != 2).  Overwrites the last adjust's
 * intermediate buffers (the next adjust re-registers the pairs). */
int ampcor_hip_stage_upsample(ampcor_hip_plan *plan, int dtype, const void *tiles, size_t count,
                              float *amp);
/* run one contraction of a plain BF16 / FP16 plan alone on host operands:
 * stage 0 the coarse one (chip C, window W), 1 the refined one (rf C, rf E);
 * chip [count][c][c] and window [count][w][w] hold the operands' 16-bit
 * patterns (bfloat16 or binary16, as the plan's form), count <= pairs; c and
 * w must be the stage's chip and window sizes.  numerators receives
 * sum_ij window[u + i][v + j] chip[i][j] [count][p][p] (p = w - c + 1) as the
 * form's tiling computes it: form 0 the plan's current tiling, 1 the 16 x 16
 * tiles (chip-row ranges summed in order), 2 the streaming form, 3 the packed
 * form (p = 32 n + 1), 4 and 5 the packed form in two and four chip-row ranges
 * (summed in order); AMPCOR_HIP_EINVAL where the form does not apply.
 * Overwrites the last adjust's intermediate buffers. */
int ampcor_hip_stage_contract(ampcor_hip_plan *plan, int stage, int form, const uint16_t *chip,
                              int c, const uint16_t *window, int w, size_t count,
                              float *numerators);
/* the FP64 window sums (sum t, sum t^2) over every chip-sized placement of
 * host FP32 windows [count][w][w], as the FP32 form's window-sum kernel
 * computes them for stage 0 (chip C, window W) or stage 1 (rf C, rf E):
 * sums [count][p][p][2], p = w - c + 1, count <= pairs; w must be the
 * stage's window size.  No counterpart in the ABI of the reference, whose
 * tests reach its tables through SumArea and Sequential::_sat
 * (tests/libampcor/sumarea-sum.cc:47-148, tests/cuda/tile-sat.cu:216-270).
 * Overwrites the last adjust's intermediate buffers. */
int ampcor_hip_stage_window_sums(ampcor_hip_plan *plan, int stage, const float *windows, int w,
                                 size_t count, double *sums);
/* the amplitudes the last adjust of an FP32 plan worked on: stage 0 the
 * detected chips and windows (|z| of complex pixels, |x| of real ones), stage
 * 1 the refined ones (the migrated tiles, upsampled by rf, then detected);
 * chips [pairs][c][c], windows [pairs][w][w], c and w the stage's sizes.  The
 * counterparts of the reference's detect and migrate steps (detect.cu:105-112,
 * migrate.cu), which its tests check directly
 * (tests/libampcor_cuda/detect.cc:202-215, migrate.cc:350-426) */
int ampcor_hip_stage_amplitudes(ampcor_hip_plan *plan, int stage, int c, int w, float *chips,
                                float *windows);
/* run the zoom stage alone on host refined surfaces [pairs][R][R] (the
 * reference drives _zoomcor the same way in its tests): the products then hold
 * its results -- refined peak, zoomed location (stage_locations), and offsets
 * (p + dp) / (rf zf) - m per axis, as for a window at (0, 0).  Overwrites the
 * products of the last adjust. */
int ampcor_hip_stage_zoom(ampcor_hip_plan *plan, const float *surfaces);
/* the zoom's search in the last adjust (or stage_zoom): the zoom takes the
 * first maximum over every cell of the zoomed grid, evaluating exactly a
 * window around each refined surface's largest sample and screening the rest
 * in FP16.  searched: pairs given the whole-grid search (their screen reached
 * the window's maximum, or it peaked on the window's edge); cells: cells
 * outside the windows evaluated exactly there; pairs: those that had any */
int ampcor_hip_zoom_stats(ampcor_hip_plan *plan, int64_t *cells, int64_t *pairs,
                          int64_t *searched);
/* the FP8 form's peak guard in the last adjust: pairs whose FP8 surface had
 * more than one placement within 0.75 / chip of its peak (each re-scored by
 * the FP64 correlation of the raster amplitudes), pairs whose peak location
 * moved, and pairs with more candidates than the guard re-scores (16; their
 * FP8 peak is kept).  All zero for the other forms. */
int ampcor_hip_fp8_guard_stats(ampcor_hip_plan *plan, int64_t *rescored, int64_t *moved,
                               int64_t *overflow);
/* per-stage device time of the last adjust, milliseconds, measured with HIP
 * events on the plan's stream; names[i] point to static strings.  Timing is
 * off until enabled, because events serialize the stream. */
int ampcor_hip_set_timing(ampcor_hip_plan *plan, int enabled);
/* times candidate tilings of both contractions on the data of the last
 * adjust (reps launches each) and keeps the fastest; the next adjust uses
 * them.  No counterpart in the reference, whose launch shapes are fixed
 * templates (lib/libampcor_cuda/correlators/correlate.cu:69-135). */
int ampcor_hip_plan_autotune(ampcor_hip_plan *plan, int reps);
int ampcor_hip_stage_times(ampcor_hip_plan *plan, int capacity, int *count, float *ms,
                           const char **names);

/* -- allocation-free adjust ----------------------------------------------------
 *
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one whole adjust -- the plan of ampcor_hip_plan_create,
 * the pairs of ampcor_hip_add_pairs and ampcor_hip_adjust -- that calls no
 * device allocator.  Its intermediate buffers and tables come from a fixed
 * scratch array of AMPCOR_HIP_NOALLOC_SCRATCH_BYTES in device memory, which
 * every device the library runs on holds until the process ends; a batch that
 * needs more is refused with AMPCOR_HIP_EINVAL (the message says how many
 * pairs of the geometry fit).  Calls on one device share that array: calls
 * that overlap on one device race.  Rasters (lines x samples, row-major,
 * dtype as for ampcor_hip_raster_create) and the four outputs (offsets
 * [pairs][2], coarse peak, refined peak, SNR [pairs]) are device pointers the
 * caller owns; the corners are host arrays as for ampcor_hip_add_pairs.  The
 * work runs on `stream` (a hipStream_t; NULL for the default stream), which
 * the call synchronizes before it returns.  Tilings are the plan's defaults
 * (no autotuning), so the results equal an untuned plan's bit for bit. */
#define AMPCOR_HIP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int ampcor_hip_adjust_noalloc(int device, void *stream, int dtype, const void *reference,
                              size_t ref_lines, size_t ref_samples, const void *target,
                              size_t tgt_lines, size_t tgt_samples, size_t pairs,
                              const int64_t *ref_begin, const int64_t *tgt_begin, size_t chip,
                              size_t window, size_t refine_margin, size_t refine_factor,
                              size_t zoom_factor, int precision, float *offsets,
                              float *coarse_peak, float *refined_peak, float *snr);

/* -- serving: continuous batching of offset-tracking jobs --------------------
 *
 * No counterpart in the reference, which correlates one plan at a time
 * (pkg/ampcor/correlators/MGA.py:52-98).  A server owns a scheduler thread.
 * Jobs are admitted under an HBM budget, queued per geometry (chip, window,
 * refine margin and factor, zoom, precision, pixel type), and their pairs are
 * batched -- across jobs and rasters -- into launches of up to batch_pairs
 * pairs on slots_per_lane concurrent plans per geometry.  With use_graphs each
 * batch replays a HIP graph captured for its size bucket.  The first batch of
 * a geometry runs the plan autotuner.  Rasters must outlive their jobs.
 * The budget bounds the server's own device memory (each lane's plans:
 * slots_per_lane x batch_pairs x the geometry's bytes per pair); rasters are
 * the caller's, resident before their jobs are submitted, so a caller gives
 * the server the HBM its rasters leave.  A lane that does not fit evicts idle
 * lanes (least recently used first) or waits for running ones to finish.
 */
typedef struct ampcor_hip_server ampcor_hip_server;

int ampcor_hip_server_create(int device, size_t hbm_budget_bytes, size_t batch_pairs,
                             int slots_per_lane, int use_graphs, ampcor_hip_server **out);
/* enqueue a job of `pairs` tile pairs (int64 (line, sample) corners); returns its id */
int ampcor_hip_server_submit(ampcor_hip_server *server, const ampcor_hip_raster *reference,
                             const ampcor_hip_raster *target, size_t pairs,
                             const int64_t *ref_begin, const int64_t *tgt_begin, size_t chip,
                             size_t window, size_t refine_margin, size_t refine_factor,
                             size_t zoom_factor, int precision, uint64_t *job);
/* the number of pairs of a submitted job (what its output buffers must hold) */
int ampcor_hip_server_job_pairs(ampcor_hip_server *server, uint64_t job, size_t *pairs);
/* wait up to timeout_ms (< 0: forever) for a job; copies its products (any
 * pointer may be NULL; each holds `capacity` pairs, at least the job's) and
 * its submit-to-finish latency, then forgets it.  AMPCOR_HIP_ESTATE if it is
 * still running at the timeout; AMPCOR_HIP_EINVAL if the buffers are too
 * small or another thread waits for the same job.  A job whose lane could not
 * be built fails alone, with that error. */
int ampcor_hip_server_wait(ampcor_hip_server *server, uint64_t job, int timeout_ms,
                           size_t capacity, float *offsets, float *coarse_peak,
                           float *refined_peak, float *snr, double *latency_ms);
int ampcor_hip_server_stats(ampcor_hip_server *server, uint64_t *pairs, uint64_t *batches,
                            uint64_t *graph_launches, uint64_t *lanes, uint64_t *tunes);
/* stops the scheduler after the batches in flight finish */
int ampcor_hip_server_destroy(ampcor_hip_server *server);

#ifdef __cplusplus
}
#endif

#endif
