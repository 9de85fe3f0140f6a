/*
 * ampcor_oracle.h -- CPU restatement of ampcor's batched tile correlator.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * CPython module, the Python strategy) links, loads or calls this code.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * there only as the checker or as the timed CPU baseline.
 *
 * Two forms of the same pipeline:
 *
 *   ORACLE_FP64  the accuracy reference.  Follows the reference algorithm in
 *                FP64: amplitude detection, zero-mean chip, SumArea window
 *                mean, direct sum of r'(t - mu) and (t - mu)^2 (the loop nest
 *                of lib/libampcor/correlators/Sequential.cc:37-84 with the
 *                formula of lib/libampcor_cuda/correlators/correlate.cu:213-225,
 *                305-309), first-max argmax, nudge, band-limited upsampling,
 *                refined correlation, Dirichlet zoom, sub-pixel fit, offsets.
 *
 *   ORACLE_FP32  the plain-FP32 twin of the HIP coarse stage: the same
 *                rounding steps, the same fixed reduction trees and the same
 *                fused-multiply-add order as the GPU kernels, so that the
 *                coarse correlation surface, its peak and the nudged window
 *                are bit-identical to the GPU's.  Its refinement chain is the
 *                FP64 one.
 *
 * Parity pinning of every function is listed in the header comment of
 * ampcor_oracle.c.
 */
#ifndef AMPCOR_ORACLE_H
#define AMPCOR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_FP64 = 0, ORACLE_FP32 = 1 };
enum { ORACLE_COMPLEX64 = 0, ORACLE_FLOAT32 = 1 };

/* geometry of one correlation plan */
typedef struct {
    int chip;          /* C: reference chip edge */
    int window;        /* W = C + 2m: target search window edge */
    int refine_factor; /* rf */
    int refine_margin; /* rm */
    int zoom_factor;   /* zf */
} oracle_geometry;

/* -- stages, exposed for the pinning tests ------------------------------- */

/* inclusive, zero-bordered summed area table S[(rows+1)*(cols+1)] of t */
void oracle_sat(const double *t, int rows, int cols, double *S);
/* sum of t over rows [r0, r1) x cols [c0, c1) from a zero-bordered SAT */
double oracle_sat_sum(const double *S, int cols, int r0, int c0, int r1, int c1);

/* FP64 normalized cross correlation of one pair:
 * ref C*C, tgt W*W (amplitudes), gamma (W-C+1)^2 */
void oracle_ncc64(const double *ref, int C, const double *tgt, int W, double *gamma);
/* unnormalized cross sum: out[u][v] = sum_ij ref[i][j] * tgt[u+i][v+j] */
void oracle_cross_sum(const double *ref, int C, const double *tgt, int W, double *out);

/* first strict maximum in row-major order; NaN never wins; all-NaN -> (0,0), peak NaN */
void oracle_argmax_f32(const float *g, int rows, int cols, int *row, int *col, float *peak);
void oracle_argmax_f64(const double *g, int rows, int cols, int *row, int *col, double *peak);

/* clamp the expanded window of edge E around the peak into the W window */
void oracle_nudge(int row, int col, int refine_margin, int window, int expanded,
                  int *top, int *left);

/* band-limited (x rf) upsampling of a complex N*N tile, interleaved re/im,
 * out is (rf N)^2 interleaved; spectrum Nyquist bin split for even N */
void oracle_upsample(const double *in, int N, int rf, double *out);
/* one entry of the 1-D upsampling kernel: refined sample p, coarse sample n */
double oracle_upsample_kernel(int N, int rf, int p, int n);

/* band-limited zoom of a real R*R surface (R odd) on the full Z = zf*R grid */
void oracle_zoom_full(const double *g, int R, int zf, double *out);
/* periodic Dirichlet kernel of an odd-size R grid at x = d / zf */
double oracle_dirichlet(int R, int zf, int d);

/* vertex abscissa of the parabola through (-1, fm), (0, f0), (1, fp), clamped
 * to [-1/2, 1/2]; 0 unless it opens downward (the sub-pixel step) */
double oracle_parabola(double fm, double f0, double fp);
/* the pipeline's zoom search on one R x R refined surface (NaN counts as 0):
 * the first maximum (p, q) of the whole Z x Z zoom, its value and the
 * sub-pixel steps of the parabolas through its neighbours; 0 on success */
int oracle_zoom_search(const double *g, int R, int zf, int *p, int *q, double *peak, double *dp,
                       double *dq);

/* coarse SNR: peak^2 / mean(gamma^2) outside a (2*exclude+1)^2 box */
double oracle_snr(const double *g, int rows, int cols, int prow, int pcol, int exclude);

/* -- the whole pipeline --------------------------------------------------- */

typedef struct {
    float *offsets;       /* [n][2] (dLine, dSample)               required */
    float *coarse_peak;   /* [n] gamma at the coarse peak           optional */
    float *refined_peak;  /* [n] zoomed peak value                  optional */
    float *snr;           /* [n]                                    optional */
    int32_t *coarse_loc;  /* [n][2] coarse argmax (row, col)        optional */
    int32_t *window_loc;  /* [n][2] nudged (top, left)              optional */
    int32_t *zoom_loc;    /* [n][2] zoomed argmax (p, q)            optional */
    float *gamma;         /* [n][P][P] coarse surface               optional */
    float *gamma_refined; /* [n][R][R] refined surface              optional */
} oracle_products;

/*
 * Correlate n pairs.  Rasters are row-major (lines, samples), complex64
 * (interleaved float re/im) or float32 amplitude.  ref_begin / tgt_begin hold
 * the upper-left corners (line, sample) of the C*C chips and W*W windows.
 * Returns 0 on success, nonzero on invalid arguments.
 */
int oracle_adjust(int form, int dtype,
                  const void *ref_raster, int64_t ref_lines, int64_t ref_samples,
                  const void *tgt_raster, int64_t tgt_lines, int64_t tgt_samples,
                  int64_t n, const int32_t *ref_begin, const int32_t *tgt_begin,
                  oracle_geometry geom, int threads, oracle_products out);

#ifdef __cplusplus
}
#endif

#endif
