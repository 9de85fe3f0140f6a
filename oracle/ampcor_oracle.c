/*
 * ampcor_oracle.c -- CPU restatement of ampcor's batched tile correlator.
 *
 * TEST INFRASTRUCTURE ONLY (see ampcor_oracle.h).  The product never links,
 * loads or calls this file; it is the checker and the timed CPU baseline.
 *
 * Reference files followed (paths relative to the ampcor source tree):
 *   detection        lib/libampcor_cuda/correlators/detect.cu:105-112 (|z|)
 *   chip statistics  lib/libampcor_cuda/correlators/refStats.cu:141-297
 *                    (mean removed in place, stats = sqrt(sum r'^2))
 *   SAT / SumArea    lib/libampcor/correlators/SumArea.icc:67-180
 *   window mean      lib/libampcor_cuda/correlators/tgtStats.cu:127-163
 *   NCC              lib/libampcor_cuda/correlators/correlate.cu:213-225,305-309
 *                    loop nest of lib/libampcor/correlators/Sequential.cc:37-84
 *   argmax           lib/libampcor_cuda/correlators/maxcor.cu:110-129
 *                    (first strict maximum; here NaN never wins)
 *   nudge            lib/libampcor_cuda/correlators/nudge.cu:113-139
 *   migrate          lib/libampcor_cuda/correlators/migrate.cu:132-151
 *   refined layouts  lib/libampcor_cuda/correlators/Sequential.icc:184-188
 *   offsets          lib/libampcor_cuda/correlators/offsets.cu:124-125,
 *                    Sequential.icc:1298-1300, with the refine margin counted
 *                    once (the reference subtracts it twice, because _nudge
 *                    rewrites the coarse peak in place, Sequential.icc:97).
 *   plan geometry    pkg/ampcor/correlators/Plan.py:160-183 (caller side)
 *
 * The reference's refinement chain (Sequential.icc:704-1286) is defective:
 * its spectrum embedding has no Nyquist split and no scaling, its target
 * "inverse" FFT runs forward, and its zoom keeps |z|.  The refinement here is
 * defined from the mathematics instead:
 *   upsampling  band-limited x rf interpolation: DFT, zero-pad with the
 *               Nyquist bin split in half for even N, inverse DFT, 1/N^2, so
 *               refined sample rf*n equals coarse sample n;
 *   zoom        band-limited x zf interpolation of the real refined surface
 *               (R = 2 rf rm + 1 is odd: periodic Dirichlet kernel) on the
 *               whole Z x Z grid, first maximum over every cell (as
 *               Sequential.icc:122-124 zooms the whole surface);
 *   sub-pixel   three-point parabola per axis through the zoomed peak (none
 *               at the edge of the zoomed grid).
 *
 * Parity pins (tests/test_oracle.py):
 *   oracle_sat / oracle_sat_sum    SumArea closed forms on ones (S = (i+1)(j+1),
 *                                  window sums = areas) + brute force;
 *   oracle_ncc64                   planted chip gamma = 1; brute-force NumPy
 *                                  NCC on small tiles; scale/offset invariance;
 *   oracle_cross_sum               tile-correlate closed form (C-|m-r|)(C-|m-c|);
 *   oracle_argmax_*                planted peaks, crafted ties and NaNs;
 *   oracle_nudge                   integer closed forms (bounds + identity);
 *   oracle_upsample(_kernel)       exact on band-limited complex exponentials,
 *                                  NumPy FFT zero-pad golden vectors, identity
 *                                  on the rf-subgrid;
 *   oracle_dirichlet/zoom_full     NumPy FFT zoom golden vectors; identity on
 *                                  the zf-subgrid; cosine bump peak location;
 *   oracle_parabola                exact vertex of sampled quadratics;
 *   oracle_zoom_search             its (p, q) and peak are the argmax of the
 *                                  NumPy FFT zoom (golden surfaces, random and
 *                                  known-shift refined surfaces, peaks more
 *                                  than a refined sample from the surface's
 *                                  argmax); off-grid known shifts recovered;
 *   oracle_snr                     closed form on a crafted surface;
 *   oracle_adjust (FP64)           Fourier-shifted speckle recovers the known
 *                                  shift; planted chip recovers the placement;
 *   oracle_adjust (FP32 twin)      agrees with the FP64 form on coarse peaks
 *                                  (and on every gamma to FP32 rounding).
 */
#include "ampcor_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------- */
/* summed area tables                                                       */
/* ---------------------------------------------------------------------- */

void oracle_sat(const double *t, int rows, int cols, double *S)
{
    const int sc = cols + 1;
    for (int j = 0; j <= cols; ++j) S[j] = 0.0;
    for (int i = 0; i < rows; ++i) {
        double run = 0.0;
        S[(size_t)(i + 1) * sc] = 0.0;
        for (int j = 0; j < cols; ++j) {
            run += t[(size_t)i * cols + j];
            S[(size_t)(i + 1) * sc + j + 1] = S[(size_t)i * sc + j + 1] + run;
        }
    }
}

double oracle_sat_sum(const double *S, int cols, int r0, int c0, int r1, int c1)
{
    const int sc = cols + 1;
    return ((S[(size_t)r1 * sc + c1] - S[(size_t)r0 * sc + c1]) - S[(size_t)r1 * sc + c0])
           + S[(size_t)r0 * sc + c0];
}

/* ---------------------------------------------------------------------- */
/* FP64 normalized cross correlation                                       */
/* ---------------------------------------------------------------------- */

/* zero-mean chip in place; returns sqrt(sum r'^2) */
static double chip_stats64(double *r, int cells)
{
    double sum = 0.0;
    for (int k = 0; k < cells; ++k) sum += r[k];
    const double mean = sum / cells;
    double ss = 0.0;
    for (int k = 0; k < cells; ++k) {
        r[k] -= mean;
        ss += r[k] * r[k];
    }
    return sqrt(ss);
}

/* gamma over all placements; ref must already be zero-mean with norm s */
static void ncc64_prepared(const double *ref, double s, int C, const double *tgt, int W,
                           double *S /* (W+1)^2 scratch */, double *gamma)
{
    const int P = W - C + 1;
    const double n = (double)C * C;
    oracle_sat(tgt, W, W, S);
    for (int u = 0; u < P; ++u) {
        for (int v = 0; v < P; ++v) {
            const double mu = oracle_sat_sum(S, W, u, v, u + C, v + C) / n;
            double num = 0.0, var = 0.0;
            for (int i = 0; i < C; ++i) {
                const double *trow = tgt + (size_t)(u + i) * W + v;
                const double *rrow = ref + (size_t)i * C;
                for (int j = 0; j < C; ++j) {
                    const double d = trow[j] - mu;
                    num += rrow[j] * d;
                    var += d * d;
                }
            }
            gamma[(size_t)u * P + v] = (var > 0.0 && s > 0.0) ? num / (s * sqrt(var)) : NAN;
        }
    }
}

void oracle_ncc64(const double *ref, int C, const double *tgt, int W, double *gamma)
{
    double *r = (double *)malloc(sizeof(double) * C * C);
    double *S = (double *)malloc(sizeof(double) * (W + 1) * (W + 1));
    memcpy(r, ref, sizeof(double) * C * C);
    const double s = chip_stats64(r, C * C);
    ncc64_prepared(r, s, C, tgt, W, S, gamma);
    free(S);
    free(r);
}

void oracle_cross_sum(const double *ref, int C, const double *tgt, int W, double *out)
{
    const int P = W - C + 1;
    for (int u = 0; u < P; ++u)
        for (int v = 0; v < P; ++v) {
            double acc = 0.0;
            for (int i = 0; i < C; ++i)
                for (int j = 0; j < C; ++j)
                    acc += ref[(size_t)i * C + j] * tgt[(size_t)(u + i) * W + v + j];
            out[(size_t)u * P + v] = acc;
        }
}

/* ---------------------------------------------------------------------- */
/* argmax and nudge                                                         */
/* ---------------------------------------------------------------------- */

void oracle_argmax_f32(const float *g, int rows, int cols, int *row, int *col, float *peak)
{
    long best = -1;
    float high = 0.0f;
    for (long k = 0; k < (long)rows * cols; ++k) {
        const float x = g[k];
        if (isnan(x)) continue;
        if (best < 0 || x > high) { high = x; best = k; }
    }
    if (best < 0) { *row = 0; *col = 0; *peak = NAN; return; }
    *row = (int)(best / cols);
    *col = (int)(best % cols);
    *peak = high;
}

void oracle_argmax_f64(const double *g, int rows, int cols, int *row, int *col, double *peak)
{
    long best = -1;
    double high = 0.0;
    for (long k = 0; k < (long)rows * cols; ++k) {
        const double x = g[k];
        if (isnan(x)) continue;
        if (best < 0 || x > high) { high = x; best = k; }
    }
    if (best < 0) { *row = 0; *col = 0; *peak = NAN; return; }
    *row = (int)(best / cols);
    *col = (int)(best % cols);
    *peak = high;
}

void oracle_nudge(int row, int col, int refine_margin, int window, int expanded,
                  int *top, int *left)
{
    int l = col - refine_margin;
    if (l < 0) l = 0;
    if (l + expanded > window) l = window - expanded;
    int t = row - refine_margin;
    if (t < 0) t = 0;
    if (t + expanded > window) t = window - expanded;
    *top = t;
    *left = l;
}

/* ---------------------------------------------------------------------- */
/* band-limited upsampling and zoom                                         */
/* ---------------------------------------------------------------------- */

double oracle_upsample_kernel(int N, int rf, int p, int n)
{
    if (p % rf == 0) return (p / rf == n) ? 1.0 : 0.0;
    /* explicit spectral sum: (1/N) sum_k w_k exp(2 pi i k t / N), t = p/rf - n */
    const double t = (double)(p - rf * n) / rf;
    const double theta = 2.0 * M_PI * t / N;
    const int K = (N % 2) ? (N - 1) / 2 : N / 2 - 1;
    double sum = 1.0;
    for (int k = 1; k <= K; ++k) sum += 2.0 * cos(k * theta);
    if (N % 2 == 0) sum += cos(0.5 * N * theta); /* split Nyquist bin: two halves */
    return sum / N;
}

void oracle_upsample(const double *in, int N, int rf, double *out)
{
    const int M = rf * N;
    double *U = (double *)malloc(sizeof(double) * M * N);
    double *T = (double *)malloc(sizeof(double) * 2 * N * M);
    for (int p = 0; p < M; ++p)
        for (int n = 0; n < N; ++n) U[(size_t)p * N + n] = oracle_upsample_kernel(N, rf, p, n);
    /* T[n][q] = sum_m x[n][m] U[q][m] */
    for (int n = 0; n < N; ++n)
        for (int q = 0; q < M; ++q) {
            double re = 0.0, im = 0.0;
            for (int m = 0; m < N; ++m) {
                const double u = U[(size_t)q * N + m];
                re += in[2 * ((size_t)n * N + m)] * u;
                im += in[2 * ((size_t)n * N + m) + 1] * u;
            }
            T[2 * ((size_t)n * M + q)] = re;
            T[2 * ((size_t)n * M + q) + 1] = im;
        }
    /* y[p][q] = sum_n U[p][n] T[n][q] */
    for (int p = 0; p < M; ++p)
        for (int q = 0; q < M; ++q) {
            double re = 0.0, im = 0.0;
            for (int n = 0; n < N; ++n) {
                const double u = U[(size_t)p * N + n];
                re += u * T[2 * ((size_t)n * M + q)];
                im += u * T[2 * ((size_t)n * M + q) + 1];
            }
            out[2 * ((size_t)p * M + q)] = re;
            out[2 * ((size_t)p * M + q) + 1] = im;
        }
    free(T);
    free(U);
}

double oracle_dirichlet(int R, int zf, int d)
{
    if (d % zf == 0) return ((d / zf) % R == 0) ? 1.0 : 0.0;
    const double x = (double)d / zf;
    return sin(M_PI * x) / (R * sin(M_PI * x / R));
}

void oracle_zoom_full(const double *g, int R, int zf, double *out)
{
    const int Z = zf * R;
    double *H = (double *)malloc(sizeof(double) * Z * R);
    for (int p = 0; p < Z; ++p)
        for (int b = 0; b < R; ++b) {
            double acc = 0.0;
            for (int a = 0; a < R; ++a) acc += oracle_dirichlet(R, zf, p - zf * a) * g[a * R + b];
            H[(size_t)p * R + b] = acc;
        }
    for (int p = 0; p < Z; ++p)
        for (int q = 0; q < Z; ++q) {
            double acc = 0.0;
            for (int b = 0; b < R; ++b) acc += H[(size_t)p * R + b] * oracle_dirichlet(R, zf, q - zf * b);
            out[(size_t)p * Z + q] = acc;
        }
    free(H);
}

double oracle_snr(const double *g, int rows, int cols, int prow, int pcol, int exclude)
{
    const double peak = g[(size_t)prow * cols + pcol];
    if (isnan(peak)) return NAN;
    double sum = 0.0;
    long count = 0;
    for (int u = 0; u < rows; ++u)
        for (int v = 0; v < cols; ++v) {
            if (abs(u - prow) <= exclude && abs(v - pcol) <= exclude) continue;
            const double x = g[(size_t)u * cols + v];
            if (isnan(x)) continue;
            sum += x * x;
            ++count;
        }
    if (count == 0 || sum <= 0.0) return INFINITY;
    return peak * peak / (sum / count);
}

/* ---------------------------------------------------------------------- */
/* the FP32 twin of the HIP coarse stage                                    */
/* ---------------------------------------------------------------------- */

/* The HIP chip-statistics kernel runs 256 threads per pair: thread t sums
 * cells t, t+256, ... in order, then a halving tree over the 256 partials. */
#define TWIN_THREADS 256
static double twin_tree(double *partial)
{
    for (int s = TWIN_THREADS / 2; s > 0; s >>= 1)
        for (int t = 0; t < s; ++t) partial[t] = partial[t] + partial[t + s];
    return partial[0];
}

/* zero-mean FP32 chip (in place) + (sum r', sqrt(sum r'^2)) exactly as the GPU */
static void twin_chip_stats(float *r, int cells, double *sum_rp, double *norm)
{
    double p1[TWIN_THREADS], p2[TWIN_THREADS];
    for (int t = 0; t < TWIN_THREADS; ++t) {
        double acc = 0.0;
        for (int k = t; k < cells; k += TWIN_THREADS) acc = acc + (double)r[k];
        p1[t] = acc;
    }
    const double mean = twin_tree(p1) / (double)cells;
    for (int k = 0; k < cells; ++k) r[k] = (float)((double)r[k] - mean);
    for (int t = 0; t < TWIN_THREADS; ++t) {
        double a1 = 0.0, a2 = 0.0;
        for (int k = t; k < cells; k += TWIN_THREADS) {
            const double d = (double)r[k];
            a1 = a1 + d;
            a2 = a2 + d * d;
        }
        p1[t] = a1;
        p2[t] = a2;
    }
    *sum_rp = twin_tree(p1);
    *norm = sqrt(twin_tree(p2));
}

/* The HIP window-sum kernel gives one 64-lane wavefront to a pair.  Lane l
 * owns columns l + 64 q.  Vertical box sums start as the sum of rows 0..C-1
 * and slide down one row per placement row, V <- (V + t(u+C-1)) - t(u-1);
 * each row of V then gets an inclusive Kogge-Stone scan across the 64 lanes
 * (offsets 1..32) per 64-column chunk, plus the running carry of the earlier
 * chunks; the window sum is the difference of two scan entries. */
static void twin_scan_row(const double *V1, const double *V2, int W, double *s1, double *s2)
{
    const int Q = (W + 63) / 64;
    double v1[64], v2[64], o1[64], o2[64];
    double carry1 = 0.0, carry2 = 0.0;
    for (int q = 0; q < Q; ++q) {
        for (int l = 0; l < 64; ++l) {
            const int x = l + 64 * q;
            v1[l] = (x < W) ? V1[x] : 0.0;
            v2[l] = (x < W) ? V2[x] : 0.0;
        }
        for (int d = 1; d < 64; d <<= 1) {
            for (int l = 0; l < 64; ++l) { o1[l] = v1[l]; o2[l] = v2[l]; }
            for (int l = d; l < 64; ++l) {
                v1[l] = o1[l - d] + o1[l];
                v2[l] = o2[l - d] + o2[l];
            }
        }
        for (int l = 0; l < 64; ++l) {
            v1[l] = carry1 + v1[l];
            v2[l] = carry2 + v2[l];
        }
        carry1 = v1[63];
        carry2 = v2[63];
        for (int l = 0; l < 64; ++l) {
            const int x = l + 64 * q;
            if (x < W) { s1[x] = v1[l]; s2[x] = v2[l]; }
        }
    }
}

static void twin_window_sums(const float *t, int W, int C, double *sum1, double *sum2)
{
    const int P = W - C + 1;
    double V1[512], V2[512], s1[512], s2[512];
    for (int x = 0; x < W; ++x) {
        V1[x] = V2[x] = 0.0;
        for (int i = 0; i < C; ++i) {
            const double d = (double)t[(size_t)i * W + x];
            V1[x] = V1[x] + d;
            V2[x] = V2[x] + d * d;
        }
    }
    for (int u = 0; u < P; ++u) {
        if (u > 0)
            for (int x = 0; x < W; ++x) {
                const double a = (double)t[(size_t)(u + C - 1) * W + x];
                const double b = (double)t[(size_t)(u - 1) * W + x];
                V1[x] = (V1[x] + a) - b;
                V2[x] = (V2[x] + a * a) - b * b;
            }
        twin_scan_row(V1, V2, W, s1, s2);
        for (int v = 0; v < P; ++v) {
            sum1[(size_t)u * P + v] = s1[v + C - 1] - (v > 0 ? s1[v - 1] : 0.0);
            sum2[(size_t)u * P + v] = s2[v + C - 1] - (v > 0 ? s2[v - 1] : 0.0);
        }
    }
}

/* coarse gamma, bit-identical to the HIP kernels' */
static void twin_coarse(const float *amp_ref /* C*C, becomes r' */, const float *amp_tgt, int C,
                        int W, double *S1, double *S2, float *rT, float *gamma)
{
    const int P = W - C + 1;
    double sum_rp, norm;
    float *r = (float *)amp_ref;
    twin_chip_stats(r, C * C, &sum_rp, &norm);
    for (int i = 0; i < C; ++i)
        for (int j = 0; j < C; ++j) rT[(size_t)j * C + i] = r[(size_t)i * C + j];
    twin_window_sums(amp_tgt, W, C, S1, S2);
    const double n = (double)C * C;
    for (int u = 0; u < P; ++u)
        for (int v = 0; v < P; ++v) {
            /* one FP32 fused-multiply-add chain per chip column (rows in
             * order, from zero); the column sums added in FP64, in order */
            double num = 0.0;
            for (int j = 0; j < C; ++j) {
                const float *rc = rT + (size_t)j * C;
                float col = 0.0f;
                for (int i = 0; i < C; ++i) col = fmaf(rc[i], amp_tgt[(size_t)(u + i) * W + v + j], col);
                num = num + (double)col;
            }
            const double s1 = S1[(size_t)u * P + v];
            const double s2 = S2[(size_t)u * P + v];
            const double mu = s1 / n;
            const double var = s2 - mu * s1;
            const double A = mu * sum_rp;
            const double B = (var > 0.0 && norm > 0.0) ? norm * sqrt(var) : 0.0;
            gamma[(size_t)u * P + v] = (B > 0.0) ? (float)((num - A) / B) : NAN;
        }
}

/* ---------------------------------------------------------------------- */
/* the pipeline                                                             */
/* ---------------------------------------------------------------------- */

static inline void load_complex(int dtype, const void *raster, int64_t samples, int64_t line,
                                int64_t sample, double *re, double *im)
{
    if (dtype == ORACLE_COMPLEX64) {
        const float *p = (const float *)raster + 2 * (line * samples + sample);
        *re = p[0];
        *im = p[1];
    } else {
        *re = ((const float *)raster)[line * samples + sample];
        *im = 0.0;
    }
}

static inline double amp64(double re, double im) { return sqrt(re * re + im * im); }

/* three-point parabola through (-1, fm), (0, f0), (+1, fp): the abscissa of
 * its vertex, clamped to [-1/2, 1/2]; 0 unless the parabola opens downward */
double oracle_parabola(double fm, double f0, double fp)
{
    const double den = fm - 2.0 * f0 + fp;
    double d = 0.0;
    if (den < 0.0) d = 0.5 * (fm - fp) / den;
    return d < -0.5 ? -0.5 : (d > 0.5 ? 0.5 : d);
}

/* full-grid zoom search, as the reference zooms the whole refined surface and
 * takes the maximum over every zoomed cell (Sequential.icc:122-124): f = the
 * Z x Z band-limited zoom (H[p][b] = sum_a D(p - zf a) g[a][b], then
 * f[p][q] = sum_b H[p][b] D(q - zf b), sums in index order), its first
 * maximum in row-major order, and the parabola through the neighbours of that
 * cell on each axis (none at the edge of the grid).  FP64. */
static void zoom_peak(const double *g /* R*R, NaN -> 0 */, int R, int zf, const double *D /* index d + Z */,
                      double *H /* Z*R */, double *f /* Z*Z */, int *pp, int *pq, double *fpeak,
                      double *dp, double *dq)
{
    const int Z = zf * R;
    for (int p = 0; p < Z; ++p)
        for (int b = 0; b < R; ++b) {
            double acc = 0.0;
            for (int a = 0; a < R; ++a) acc += D[p - zf * a + Z] * g[a * R + b];
            H[(size_t)p * R + b] = acc;
        }
    for (int p = 0; p < Z; ++p)
        for (int q = 0; q < Z; ++q) {
            double acc = 0.0;
            for (int b = 0; b < R; ++b) acc += H[(size_t)p * R + b] * D[q - zf * b + Z];
            f[(size_t)p * Z + q] = acc;
        }
    size_t bi = 0;
    for (size_t k = 1; k < (size_t)Z * Z; ++k)
        if (f[k] > f[bi]) bi = k;
    const int ip = (int)(bi / Z), iq = (int)(bi % Z);
    *pp = ip;
    *pq = iq;
    *fpeak = f[bi];
    *dp = (ip > 0 && ip < Z - 1) ? oracle_parabola(f[bi - Z], f[bi], f[bi + Z]) : 0.0;
    *dq = (iq > 0 && iq < Z - 1) ? oracle_parabola(f[bi - 1], f[bi], f[bi + 1]) : 0.0;
}

int oracle_zoom_search(const double *g, int R, int zf, int *p, int *q, double *peak, double *dp,
                       double *dq)
{
    if (R < 1 || zf < 1) return 1;
    const int Z = zf * R;
    double *D = (double *)malloc(sizeof(double) * (2 * Z + 1));
    double *h = (double *)malloc(sizeof(double) * R * R);
    double *H = (double *)malloc(sizeof(double) * Z * R);
    double *f = (double *)malloc(sizeof(double) * Z * Z);
    for (int d = -Z; d <= Z; ++d) D[d + Z] = oracle_dirichlet(R, zf, d);
    for (int c = 0; c < R * R; ++c) h[c] = isnan(g[c]) ? 0.0 : g[c];
    zoom_peak(h, R, zf, D, H, f, p, q, peak, dp, dq);
    free(f);
    free(H);
    free(h);
    free(D);
    return 0;
}

int oracle_adjust(int form, int dtype,
                  const void *ref_raster, int64_t ref_lines, int64_t ref_samples,
                  const void *tgt_raster, int64_t tgt_lines, int64_t tgt_samples,
                  int64_t n, const int32_t *ref_begin, const int32_t *tgt_begin,
                  oracle_geometry g, int threads, oracle_products out)
{
    const int C = g.chip, W = g.window, rf = g.refine_factor, rm = g.refine_margin, zf = g.zoom_factor;
    if (C < 2 || W < C || (W - C) % 2 || rf < 1 || rm < 0 || zf < 1) return 1;
    const int m = (W - C) / 2;
    if (rm > m) return 1;
    if (form != ORACLE_FP64 && form != ORACLE_FP32) return 1;
    if (dtype != ORACLE_COMPLEX64 && dtype != ORACLE_FLOAT32) return 1;
    if (!out.offsets) return 1;
    const int P = W - C + 1, E = C + 2 * rm, Cr = rf * C, Wr = rf * E, R = Wr - Cr + 1, Z = zf * R;
    if (R > 129) return 1;
    for (int64_t k = 0; k < n; ++k) {
        if (ref_begin[2 * k] < 0 || ref_begin[2 * k + 1] < 0 || ref_begin[2 * k] + C > ref_lines ||
            ref_begin[2 * k + 1] + C > ref_samples)
            return 2;
        if (tgt_begin[2 * k] < 0 || tgt_begin[2 * k + 1] < 0 || tgt_begin[2 * k] + W > tgt_lines ||
            tgt_begin[2 * k + 1] + W > tgt_samples)
            return 2;
    }
    /* Dirichlet table for the zoom, index d + Z for d in [-Z, Z] */
    double *D = (double *)malloc(sizeof(double) * (2 * Z + 1));
    for (int d = -Z; d <= Z; ++d) D[d + Z] = oracle_dirichlet(R, zf, d);

#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
#endif
    {
        double *ref64 = (double *)malloc(sizeof(double) * C * C);
        double *tgt64 = (double *)malloc(sizeof(double) * W * W);
        const int Wmax = (W > Wr ? W : Wr) + 1;
        double *S = (double *)malloc(sizeof(double) * Wmax * Wmax * 2);
        double *g64 = (double *)malloc(sizeof(double) * P * P);
        float *ref32 = (float *)malloc(sizeof(float) * C * C);
        float *tgt32 = (float *)malloc(sizeof(float) * W * W);
        float *rT = (float *)malloc(sizeof(float) * C * C);
        float *g32 = (float *)malloc(sizeof(float) * P * P);
        double *cref = (double *)malloc(sizeof(double) * 2 * C * C);
        double *ctgt = (double *)malloc(sizeof(double) * 2 * E * E);
        double *uref = (double *)malloc(sizeof(double) * 2 * Cr * Cr);
        double *utgt = (double *)malloc(sizeof(double) * 2 * Wr * Wr);
        double *aref = (double *)malloc(sizeof(double) * Cr * Cr);
        double *atgt = (double *)malloc(sizeof(double) * Wr * Wr);
        double *gR = (double *)malloc(sizeof(double) * R * R);
        double *zH = (double *)malloc(sizeof(double) * Z * R);
        double *zf_ = (double *)malloc(sizeof(double) * Z * Z);

#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t k = 0; k < n; ++k) {
            const int64_t rl = ref_begin[2 * k], rs = ref_begin[2 * k + 1];
            const int64_t tl = tgt_begin[2 * k], ts = tgt_begin[2 * k + 1];
            int row, col;
            double peak;
            /* ---- coarse ---- */
            for (int i = 0; i < C; ++i)
                for (int j = 0; j < C; ++j) {
                    double re, im;
                    load_complex(dtype, ref_raster, ref_samples, rl + i, rs + j, &re, &im);
                    const double a = dtype == ORACLE_COMPLEX64 ? amp64(re, im) : fabs(re);
                    ref64[i * C + j] = a;
                    ref32[i * C + j] = (float)a;
                }
            for (int i = 0; i < W; ++i)
                for (int j = 0; j < W; ++j) {
                    double re, im;
                    load_complex(dtype, tgt_raster, tgt_samples, tl + i, ts + j, &re, &im);
                    const double a = dtype == ORACLE_COMPLEX64 ? amp64(re, im) : fabs(re);
                    tgt64[i * W + j] = a;
                    tgt32[i * W + j] = (float)a;
                }
            if (form == ORACLE_FP64) {
                const double s = chip_stats64(ref64, C * C);
                ncc64_prepared(ref64, s, C, tgt64, W, S, g64);
                oracle_argmax_f64(g64, P, P, &row, &col, &peak);
                if (out.gamma)
                    for (int c = 0; c < P * P; ++c) out.gamma[k * P * P + c] = (float)g64[c];
            } else {
                twin_coarse(ref32, tgt32, C, W, S, S + (size_t)(W + 1) * (W + 1), rT, g32);
                float pk;
                oracle_argmax_f32(g32, P, P, &row, &col, &pk);
                peak = pk;
                for (int c = 0; c < P * P; ++c) g64[c] = g32[c];
                if (out.gamma) memcpy(out.gamma + k * P * P, g32, sizeof(float) * P * P);
            }
            const double snr = oracle_snr(g64, P, P, row, col, rf);
            int top, left;
            oracle_nudge(row, col, rm, W, E, &top, &left);

            /* ---- refinement (FP64 in both forms) ---- */
            for (int i = 0; i < C; ++i)
                for (int j = 0; j < C; ++j)
                    load_complex(dtype, ref_raster, ref_samples, rl + i, rs + j,
                                 &cref[2 * (i * C + j)], &cref[2 * (i * C + j) + 1]);
            for (int i = 0; i < E; ++i)
                for (int j = 0; j < E; ++j)
                    load_complex(dtype, tgt_raster, tgt_samples, tl + top + i, ts + left + j,
                                 &ctgt[2 * (i * E + j)], &ctgt[2 * (i * E + j) + 1]);
            oracle_upsample(cref, C, rf, uref);
            oracle_upsample(ctgt, E, rf, utgt);
            for (int c = 0; c < Cr * Cr; ++c) aref[c] = amp64(uref[2 * c], uref[2 * c + 1]);
            for (int c = 0; c < Wr * Wr; ++c) atgt[c] = amp64(utgt[2 * c], utgt[2 * c + 1]);
            {
                const double s = chip_stats64(aref, Cr * Cr);
                ncc64_prepared(aref, s, Cr, atgt, Wr, S, gR);
            }
            if (out.gamma_refined)
                for (int c = 0; c < R * R; ++c) out.gamma_refined[k * R * R + c] = (float)gR[c];
            int a, b;
            double rpeak;
            oracle_argmax_f64(gR, R, R, &a, &b, &rpeak);
            for (int c = 0; c < R * R; ++c)
                if (isnan(gR[c])) gR[c] = 0.0;
            int pp, pq;
            double fpeak, dp, dq;
            zoom_peak(gR, R, zf, D, zH, zf_, &pp, &pq, &fpeak, &dp, &dq);
            const double zoom = (double)rf * zf;
            float dl = (float)((double)(top - m) + ((double)pp + dp) / zoom);
            float ds = (float)((double)(left - m) + ((double)pq + dq) / zoom);
            if (isnan(peak) || isnan(rpeak)) {
                dl = ds = NAN;
                fpeak = NAN;
            }
            out.offsets[2 * k] = dl;
            out.offsets[2 * k + 1] = ds;
            if (out.coarse_peak) out.coarse_peak[k] = (float)peak;
            if (out.refined_peak) out.refined_peak[k] = (float)fpeak;
            if (out.snr) out.snr[k] = (float)snr;
            if (out.coarse_loc) { out.coarse_loc[2 * k] = row; out.coarse_loc[2 * k + 1] = col; }
            if (out.window_loc) { out.window_loc[2 * k] = top; out.window_loc[2 * k + 1] = left; }
            if (out.zoom_loc) { out.zoom_loc[2 * k] = pp; out.zoom_loc[2 * k + 1] = pq; }
        }
        free(zf_); free(zH);
        free(gR); free(atgt); free(aref); free(utgt); free(uref); free(ctgt); free(cref);
        free(g32); free(rT); free(tgt32); free(ref32); free(g64); free(S); free(tgt64); free(ref64);
    }
    free(D);
    return 0;
}
