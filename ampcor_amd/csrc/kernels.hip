// kernels.hip -- gfx950 kernels of the batched tile correlator.
//
// The coarse stage (gather/detect, chip statistics, SAT, contraction,
// normalization, peak) rounds exactly where the CPU oracle's FP32 twin does, so
// its correlation surfaces are bit-identical: FP64 sums run in fixed orders
// (a 256-way strided pass plus a halving tree for chips, a per-lane chunked
// Kogge-Stone scan for SATs), the contraction is a fused multiply-add chain in
// a fixed (chip column outer, chip row inner) order, and nothing is contracted
// implicitly (the library is built with -ffp-contract=off).
//
// Reference stages replaced (paths relative to the ampcor tree):
//   detect.cu:78-115, refStats.cu:115-302, sat.cu:81-149, tgtStats.cu:88-168,
//   correlate.cu:160-318, maxcor.cu:81-133, nudge.cu:83-143, migrate.cu:103-154,
//   the cuFFT refinement of Sequential.icc:704-1087, r2c.cu / c2r.cu and the
//   zoom of Sequential.icc:1102-1286, offsets.cu:88-129.
#include "kernels.h"

#include <math.h>

namespace ampcor_hip {

namespace {

constexpr int kWave = 64;
constexpr int kMaxGridY = 65535;

__device__ inline float detect(int dtype, const void *base, size_t offset)
{
    if (dtype == 0) {
        const float2 z = static_cast<const float2 *>(base)[offset];
        const double re = z.x, im = z.y;
        return (float)sqrt(re * re + im * im);
    }
    return fabsf(static_cast<const float *>(base)[offset]);
}

// ---------------------------------------------------------------------------
// gather + detect: chips and windows straight from the resident rasters
// (replaces addReferenceTile/addTargetTile + _push + _detect)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather_detect(const PairDesc *__restrict__ desc, int dtype,
                                                       int C, int W, float *__restrict__ ref_amp,
                                                       float *__restrict__ tgt_amp)
{
    const int pair = blockIdx.y;
    const PairDesc d = desc[pair];
    const int nref = C * C;
    const int ntot = nref + W * W;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ntot; k += gridDim.x * blockDim.x) {
        if (k < nref) {
            const int i = k / C, j = k - i * C;
            const size_t off = (size_t)(d.ref_line + i) * d.ref_stride + d.ref_sample + j;
            ref_amp[(size_t)pair * nref + k] = detect(dtype, d.ref, off);
        } else {
            const int kk = k - nref;
            const int i = kk / W, j = kk - i * W;
            const size_t off = (size_t)(d.tgt_line + i) * d.tgt_stride + d.tgt_sample + j;
            tgt_amp[(size_t)pair * W * W + kk] = detect(dtype, d.tgt, off);
        }
    }
}

// ---------------------------------------------------------------------------
// chip statistics: mean removal, sum r', sqrt(sum r'^2); writes the zero-mean
// chip transposed (column-major) for the contraction's scalar reads
// ---------------------------------------------------------------------------
constexpr int kStatThreads = 256;

__device__ inline double tree256(double *p, double x)
{
    const int t = threadIdx.x;
    p[t] = x;
    __syncthreads();
    for (int s = kStatThreads / 2; s > 0; s >>= 1) {
        if (t < s) p[t] = p[t] + p[t + s];
        __syncthreads();
    }
    const double r = p[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kStatThreads) void k_chip_stats(const float *__restrict__ amp, int C,
                                                             float *__restrict__ rT,
                                                             double2 *__restrict__ stats)
{
    __shared__ double p1[kStatThreads];
    __shared__ double p2[kStatThreads];
    const int pair = blockIdx.x;
    const int t = threadIdx.x;
    const int cells = C * C;
    const float *a = amp + (size_t)pair * cells;
    float *out = rT + (size_t)pair * cells;

    double acc = 0.0;
    for (int k = t; k < cells; k += kStatThreads) acc = acc + (double)a[k];
    const double mean = tree256(p1, acc) / (double)cells;

    double a1 = 0.0, a2 = 0.0;
    for (int k = t; k < cells; k += kStatThreads) {
        const float r = (float)((double)a[k] - mean);
        const int i = k / C, j = k - i * C;
        out[(size_t)j * C + i] = r;
        const double dr = (double)r;
        a1 = a1 + dr;
        a2 = a2 + dr * dr;
    }
    // both trees share the barrier sequence; p2 is a second buffer
    p2[t] = a2;
    const double s1 = tree256(p1, a1);
    for (int s = kStatThreads / 2; s > 0; s >>= 1) {
        if (t < s) p2[t] = p2[t] + p2[t + s];
        __syncthreads();
    }
    if (t == 0) stats[pair] = make_double2(s1, sqrt(p2[0]));
}

// ---------------------------------------------------------------------------
// FP64 summed-area tables of t and t^2, zero-bordered (W+1)^2, one wavefront
// per pair: each row is scanned across lanes (Kogge-Stone, 64-column chunks
// with a running carry), and each lane accumulates its columns downward
// ---------------------------------------------------------------------------
template <int Q>
__global__ __launch_bounds__(256) void k_sat(const float *__restrict__ amp, int W, int pairs,
                                             double *__restrict__ S1, double *__restrict__ S2)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= pairs) return;
    const int sc = W + 1;
    const size_t ss = (size_t)sc * sc;
    const float *t = amp + (size_t)pair * W * W;
    double *s1 = S1 + pair * ss;
    double *s2 = S2 + pair * ss;
    for (int x = lane; x < sc; x += kWave) {
        s1[x] = 0.0;
        s2[x] = 0.0;
    }
    double c1[Q], c2[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) c1[q] = c2[q] = 0.0;

    for (int k = 0; k < W; ++k) {
        if (lane == 0) {
            s1[(size_t)(k + 1) * sc] = 0.0;
            s2[(size_t)(k + 1) * sc] = 0.0;
        }
        double carry1 = 0.0, carry2 = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int x = lane + kWave * q;
            double v1 = (x < W) ? (double)t[(size_t)k * W + x] : 0.0;
            double v2 = v1 * v1;
#pragma unroll
            for (int d = 1; d < kWave; d <<= 1) {
                const double o1 = __shfl_up(v1, d, kWave);
                const double o2 = __shfl_up(v2, d, kWave);
                if (lane >= d) {
                    v1 = o1 + v1;
                    v2 = o2 + v2;
                }
            }
            v1 = carry1 + v1;
            v2 = carry2 + v2;
            carry1 = __shfl(v1, kWave - 1, kWave);
            carry2 = __shfl(v2, kWave - 1, kWave);
            c1[q] = c1[q] + v1;
            c2[q] = c2[q] + v2;
            if (x < W) {
                s1[(size_t)(k + 1) * sc + x + 1] = c1[q];
                s2[(size_t)(k + 1) * sc + x + 1] = c2[q];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// the contraction: num(u, v) = sum_j sum_i r'(i, j) t(u + i, v + j)
//
// One workgroup per (pair, placement tile, chip-column range).  The target
// rows/columns the tile needs are staged in LDS; the zero-mean chip is read
// with wave-uniform addresses (scalar loads).  Each lane owns BU consecutive
// placement rows at one placement column and walks the chip column by column;
// down a column it slides a BU-deep register window over the target rows, so
// every LDS read feeds BU fused multiply-adds.
// ---------------------------------------------------------------------------
template <int BU>
__global__ __launch_bounds__(1024) void k_correlate(const float *__restrict__ rT,
                                                    const float *__restrict__ amp, int C, int W,
                                                    int P, int ub_extent, int vb_extent,
                                                    int tiles_u, int tiles_v, int nsplit,
                                                    float *__restrict__ partial)
{
    extern __shared__ float tile[];
    int b = blockIdx.x;
    const int part = b % nsplit;
    b /= nsplit;
    const int tv = b % tiles_v;
    b /= tiles_v;
    const int tu = b % tiles_u;
    const int pair = b / tiles_u;

    const int u0 = tu * ub_extent, v0 = tv * vb_extent;
    const int ucount = min(ub_extent, P - u0);
    const int vcount = min(vb_extent, P - v0);
    const int nub = (ucount + BU - 1) / BU;
    const int jlen = C / nsplit;
    const int j0 = part * jlen;
    const int rows = nub * BU + C - 1;
    const int cols = vcount + jlen - 1;

    // stage target rows [u0, u0 + rows) x columns [v0 + j0, v0 + j0 + cols)
    const float *src = amp + (size_t)pair * W * W;
    for (int k = threadIdx.x; k < rows * cols; k += blockDim.x) {
        const int r = k / cols, c = k - r * cols;
        const int gr = u0 + r;
        tile[k] = (gr < W) ? src[(size_t)gr * W + v0 + j0 + c] : 0.0f;
    }
    __syncthreads();

    const float *rp = rT + (size_t)pair * C * C;
    const int tasks = vcount * nub;
    for (int task = threadIdx.x; task < tasks; task += blockDim.x) {
        const int v = task % vcount;
        const int ubk = task / vcount;
        float acc[BU];
#pragma unroll
        for (int a = 0; a < BU; ++a) acc[a] = 0.0f;
        for (int jj = 0; jj < jlen; ++jj) {
            const float *rc = rp + (size_t)(j0 + jj) * C;
            const float *tc = tile + (size_t)(ubk * BU) * cols + v + jj;
            float w[BU];
#pragma unroll
            for (int a = 0; a < BU - 1; ++a) w[a] = tc[a * cols];
            for (int i0 = 0; i0 < C; i0 += BU) {
#pragma unroll
                for (int s = 0; s < BU; ++s) {
                    w[(s + BU - 1) % BU] = tc[(i0 + s + BU - 1) * cols];
                    const float r = rc[i0 + s];
#pragma unroll
                    for (int a = 0; a < BU; ++a) acc[a] = fmaf(r, w[(s + a) % BU], acc[a]);
                }
            }
        }
        float *dst = partial + ((size_t)pair * nsplit + part) * P * P;
#pragma unroll
        for (int a = 0; a < BU; ++a) {
            const int uu = ubk * BU + a;
            if (uu < ucount) dst[(size_t)(u0 + uu) * P + v0 + v] = acc[a];
        }
    }
}

// ---------------------------------------------------------------------------
// normalization: gamma = (num - mu_t sum r') / (|r'| sqrt(sum t^2 - mu_t sum t))
// from the FP64 SATs; NaN where the window or the chip has no variance
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_normalize(const float *__restrict__ partial, int nsplit,
                                                   const double *__restrict__ S1,
                                                   const double *__restrict__ S2,
                                                   const double2 *__restrict__ stats, int C,
                                                   int W, int P, int pairs,
                                                   float *__restrict__ gamma)
{
    const size_t total = (size_t)pairs * P * P;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int pp = P * P;
    const int pair = (int)(idx / pp);
    const int c = (int)(idx - (size_t)pair * pp);
    const int u = c / P, v = c - u * P;

    // fixed pairwise tree over the chip-column partials
    float part[16];
    for (int k = 0; k < nsplit; ++k) part[k] = partial[((size_t)pair * nsplit + k) * pp + c];
    for (int w = nsplit; w > 1; w >>= 1)
        for (int k = 0; k < w / 2; ++k) part[k] = part[2 * k] + part[2 * k + 1];
    const float num = part[0];

    const int sc = W + 1;
    const size_t base = (size_t)pair * sc * sc;
    const double *a1 = S1 + base, *a2 = S2 + base;
    const size_t tl = (size_t)u * sc + v, tr = (size_t)u * sc + v + C;
    const size_t bl = (size_t)(u + C) * sc + v, br = (size_t)(u + C) * sc + v + C;
    const double s1 = ((a1[br] - a1[tr]) - a1[bl]) + a1[tl];
    const double s2 = ((a2[br] - a2[tr]) - a2[bl]) + a2[tl];
    const double n = (double)C * C;
    const double mu = s1 / n;
    const double var = s2 - mu * s1;
    const double2 st = stats[pair];
    const double A = mu * st.x;
    const double B = (var > 0.0 && st.y > 0.0) ? st.y * sqrt(var) : 0.0;
    gamma[idx] = (B > 0.0) ? (float)(((double)num - A) / B) : __builtin_nanf("");
}

// ---------------------------------------------------------------------------
// coarse peak (first maximum, NaN never wins), SNR and nudge; one wave a pair
// ---------------------------------------------------------------------------
__device__ inline void better(float &best, int &bidx, float x, int k)
{
    if (x > best || (x == best && k < bidx)) {
        best = x;
        bidx = k;
    }
}

__device__ inline void wave_argmax(float &best, int &bidx)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const float ob = __shfl_xor(best, d, kWave);
        const int oi = __shfl_xor(bidx, d, kWave);
        better(best, bidx, ob, oi);
    }
}

__global__ __launch_bounds__(256) void k_coarse_peak(const float *__restrict__ gamma, int pairs,
                                                     int P, int W, int E, int rm, int exclude,
                                                     int2 *__restrict__ loc,
                                                     int2 *__restrict__ window,
                                                     float *__restrict__ peak,
                                                     float *__restrict__ snr)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= pairs) return;
    const int cells = P * P;
    const float *g = gamma + (size_t)pair * cells;
    float best = -INFINITY;
    int bidx = 0x7fffffff;
    for (int k = lane; k < cells; k += kWave) {
        const float x = g[k];
        if (x > best) {
            best = x;
            bidx = k;
        }
    }
    wave_argmax(best, bidx);
    const bool none = (bidx == 0x7fffffff);
    const int row = none ? 0 : bidx / P;
    const int col = none ? 0 : bidx - row * P;
    const float pk = none ? __builtin_nanf("") : best;

    double sum = 0.0;
    int count = 0;
    for (int k = lane; k < cells; k += kWave) {
        const int u = k / P, v = k - u * P;
        if (abs(u - row) <= exclude && abs(v - col) <= exclude) continue;
        const float x = g[k];
        if (x != x) continue;
        sum += (double)x * (double)x;
        ++count;
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        sum += __shfl_xor(sum, d, kWave);
        count += __shfl_xor(count, d, kWave);
    }
    if (lane == 0) {
        int left = col - rm;
        if (left < 0) left = 0;
        if (left + E > W) left = W - E;
        int top = row - rm;
        if (top < 0) top = 0;
        if (top + E > W) top = W - E;
        loc[pair] = make_int2(row, col);
        window[pair] = make_int2(top, left);
        peak[pair] = pk;
        const double p = (double)pk;
        float s;
        if (none) s = __builtin_nanf("");
        else if (count == 0 || sum <= 0.0) s = INFINITY;
        else s = (float)(p * p / (sum / count));
        snr[pair] = s;
    }
}

// ---------------------------------------------------------------------------
// band-limited upsampling y = U x U^T of complex N x N tiles (U real,
// M = rf N rows; rows rf*n are the identity and are copied)
// pass 1: T = x U^T, a workgroup per (pair, 8 tile rows), read from the raster
// pass 2: y = U T, a workgroup per (pair, 32 columns), detected on the way out
// ---------------------------------------------------------------------------
constexpr int kUpRows = 8;
constexpr int kUpCols = 32;

__global__ __launch_bounds__(256) void k_upsample_rows(const PairDesc *__restrict__ desc,
                                                       int dtype, int role,
                                                       const int2 *__restrict__ window, int N,
                                                       int rf, const float *__restrict__ Ut,
                                                       float2 *__restrict__ work)
{
    extern __shared__ float2 xs[];  // kUpRows x N
    const int pair = blockIdx.y;
    const int n0 = blockIdx.x * kUpRows;
    const int M = rf * N;
    const PairDesc d = desc[pair];
    const void *base = role == 0 ? d.ref : d.tgt;
    const int stride = role == 0 ? d.ref_stride : d.tgt_stride;
    int line = role == 0 ? d.ref_line : d.tgt_line;
    int sample = role == 0 ? d.ref_sample : d.tgt_sample;
    if (role == 1) {
        const int2 tl = window[pair];
        line += tl.x;
        sample += tl.y;
    }
    const int nrows = min(kUpRows, N - n0);
    for (int k = threadIdx.x; k < nrows * N; k += blockDim.x) {
        const int r = k / N, c = k - r * N;
        const size_t off = (size_t)(line + n0 + r) * stride + sample + c;
        float2 z;
        if (dtype == 0) z = static_cast<const float2 *>(base)[off];
        else z = make_float2(static_cast<const float *>(base)[off], 0.0f);
        xs[k] = z;
    }
    __syncthreads();
    float2 *out = work + (size_t)pair * N * M;
    for (int k = threadIdx.x; k < nrows * M; k += blockDim.x) {
        const int r = k / M, q = k - r * M;
        float2 acc;
        if (q % rf == 0) {
            acc = xs[r * N + q / rf];
        } else {
            float re = 0.0f, im = 0.0f;
            for (int m = 0; m < N; ++m) {
                const float u = Ut[(size_t)m * M + q];
                const float2 x = xs[r * N + m];
                re = fmaf(x.x, u, re);
                im = fmaf(x.y, u, im);
            }
            acc = make_float2(re, im);
        }
        out[(size_t)(n0 + r) * M + q] = acc;
    }
}

__global__ __launch_bounds__(256) void k_upsample_cols(const float2 *__restrict__ work, int N,
                                                       int rf, const float *__restrict__ U,
                                                       float *__restrict__ amp)
{
    extern __shared__ float2 ts[];  // N x kUpCols
    const int pair = blockIdx.y;
    const int M = rf * N;
    const int q0 = blockIdx.x * kUpCols;
    const int ncols = min(kUpCols, M - q0);
    const float2 *in = work + (size_t)pair * N * M;
    for (int k = threadIdx.x; k < N * kUpCols; k += blockDim.x) {
        const int n = k / kUpCols, c = k - n * kUpCols;
        ts[k] = (c < ncols) ? in[(size_t)n * M + q0 + c] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    float *out = amp + (size_t)pair * M * M;
    for (int k = threadIdx.x; k < M * kUpCols; k += blockDim.x) {
        const int p = k / kUpCols, c = k - p * kUpCols;
        if (c >= ncols) continue;
        float2 y;
        if (p % rf == 0) {
            y = ts[(p / rf) * kUpCols + c];
        } else {
            const float *urow = U + (size_t)p * N;
            float re = 0.0f, im = 0.0f;
            for (int n = 0; n < N; ++n) {
                const float u = urow[n];
                const float2 t = ts[n * kUpCols + c];
                re = fmaf(u, t.x, re);
                im = fmaf(u, t.y, im);
            }
            y = make_float2(re, im);
        }
        const double re = y.x, im = y.y;
        out[(size_t)p * M + q0 + c] = (float)sqrt(re * re + im * im);
    }
}

// ---------------------------------------------------------------------------
// zoom + sub-pixel peak + offsets: band-limited (Dirichlet) interpolation of
// the refined surface on the (2 zf + 1)^2 zoom cells around its peak, first
// maximum, three-point parabola per axis; one workgroup per pair
// ---------------------------------------------------------------------------
constexpr int kZoomThreads = 256;

__global__ __launch_bounds__(kZoomThreads) void k_zoom(
    const float *__restrict__ gamma_refined, int R, int zf, const double *__restrict__ D,
    const int2 *__restrict__ window, const float *__restrict__ coarse_peak, int margin, int rf,
    float *__restrict__ offsets, float *__restrict__ refined_peak, int2 *__restrict__ zoom_loc)
{
    extern __shared__ double zs[];
    const int pair = blockIdx.x;
    const int t = threadIdx.x;
    const int Z = zf * R;
    const int span = 2 * zf + 1;
    double *g = zs;                   // R*R
    double *H = g + R * R;            // span*R
    double *f = H + span * R;         // span*span
    __shared__ float rb[kZoomThreads];
    __shared__ int ri[kZoomThreads];

    const float *gr = gamma_refined + (size_t)pair * R * R;
    float best = -INFINITY;
    int bidx = 0x7fffffff;
    for (int k = t; k < R * R; k += kZoomThreads) {
        const float x = gr[k];
        g[k] = (x != x) ? 0.0 : (double)x;
        if (x > best) {
            best = x;
            bidx = k;
        }
    }
    rb[t] = best;
    ri[t] = bidx;
    __syncthreads();
    for (int s = kZoomThreads / 2; s > 0; s >>= 1) {
        if (t < s) {
            float b0 = rb[t];
            int i0 = ri[t];
            better(b0, i0, rb[t + s], ri[t + s]);
            rb[t] = b0;
            ri[t] = i0;
        }
        __syncthreads();
    }
    const bool none = (ri[0] == 0x7fffffff);
    const int a = none ? 0 : ri[0] / R;
    const int bcol = none ? 0 : ri[0] - a * R;

    const int p0 = max(zf * a - zf, 0), p1 = min(zf * a + zf, Z - 1);
    const int q0 = max(zf * bcol - zf, 0), q1 = min(zf * bcol + zf, Z - 1);
    const int np = p1 - p0 + 1, nq = q1 - q0 + 1;
    for (int k = t; k < np * R; k += kZoomThreads) {
        const int ip = k / R, bb = k - ip * R;
        double acc = 0.0;
        for (int aa = 0; aa < R; ++aa) acc += D[(p0 + ip) - zf * aa + Z] * g[aa * R + bb];
        H[k] = acc;
    }
    __syncthreads();
    for (int k = t; k < np * nq; k += kZoomThreads) {
        const int ip = k / nq, iq = k - ip * nq;
        double acc = 0.0;
        for (int bb = 0; bb < R; ++bb) acc += H[ip * R + bb] * D[(q0 + iq) - zf * bb + Z];
        f[k] = acc;
    }
    __syncthreads();
    if (t != 0) return;
    int bi = 0;
    for (int k = 1; k < np * nq; ++k)
        if (f[k] > f[bi]) bi = k;
    const int ip = bi / nq, iq = bi - ip * nq;
    double dp = 0.0, dq = 0.0;
    if (ip > 0 && ip < np - 1) {
        const double fm = f[(ip - 1) * nq + iq], f0 = f[bi], fp = f[(ip + 1) * nq + iq];
        const double den = fm - 2.0 * f0 + fp;
        if (den < 0.0) dp = 0.5 * (fm - fp) / den;
    }
    if (iq > 0 && iq < nq - 1) {
        const double fm = f[ip * nq + iq - 1], f0 = f[bi], fp = f[ip * nq + iq + 1];
        const double den = fm - 2.0 * f0 + fp;
        if (den < 0.0) dq = 0.5 * (fm - fp) / den;
    }
    dp = dp < -0.5 ? -0.5 : (dp > 0.5 ? 0.5 : dp);
    dq = dq < -0.5 ? -0.5 : (dq > 0.5 ? 0.5 : dq);
    const int pp = p0 + ip, pq = q0 + iq;
    const int2 tl = window[pair];
    const double zoom = (double)rf * zf;
    float dl = (float)((double)(tl.x - margin) + ((double)pp + dp) / zoom);
    float ds = (float)((double)(tl.y - margin) + ((double)pq + dq) / zoom);
    float fpk = (float)f[bi];
    const float cp = coarse_peak[pair];
    if (none || cp != cp) {
        dl = ds = fpk = __builtin_nanf("");
    }
    offsets[2 * pair] = dl;
    offsets[2 * pair + 1] = ds;
    refined_peak[pair] = fpk;
    zoom_loc[pair] = make_int2(pp, pq);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static size_t correlate_lds(int C, int bu, int ucount, int vcount, int nsplit)
{
    const int nub = (ucount + bu - 1) / bu;
    return (size_t)(nub * bu + C - 1) * (vcount + C / nsplit - 1) * sizeof(float);
}

CorrelateShape choose_correlate_shape(int C, int P, bool exact)
{
    CorrelateShape s{};
    s.bu = (C % 8 == 0) ? 8 : (C % 4 == 0) ? 4 : (C % 2 == 0) ? 2 : 1;
    const size_t budget = 96 * 1024;
    // chip-column split: the exact form keeps the whole contraction in a lane
    s.nsplit = 1;
    if (!exact) {
        for (int ns : {8, 4, 2})
            if (C % ns == 0 && C / ns >= 16) {
                s.nsplit = ns;
                break;
            }
    }
    int extent = P;
    const int candidates[] = {P, 64, 48, 32, 24, 16, 8};
    for (int e : candidates) {
        if (e > P) continue;
        extent = e;
        if (correlate_lds(C, s.bu, e, e, s.nsplit) <= budget) break;
    }
    s.ub = s.vb = extent;
    s.tiles_u = s.tiles_v = (P + extent - 1) / extent;
    s.lds_bytes = correlate_lds(C, s.bu, extent, extent, s.nsplit);
    const int tasks = extent * ((extent + s.bu - 1) / s.bu);
    int threads = ((tasks + kWave - 1) / kWave) * kWave;
    s.threads = threads > 1024 ? 1024 : (threads < kWave ? kWave : threads);
    return s;
}

hipError_t launch_gather_detect(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int C,
                                int W, float *ref_amp, float *tgt_amp)
{
    const int cells = C * C + W * W;
    const int bx = (cells + 256 * 8 - 1) / (256 * 8);
    for (int first = 0; first < pairs; first += kMaxGridY) {
        const int n = pairs - first < kMaxGridY ? pairs - first : kMaxGridY;
        hipLaunchKernelGGL(k_gather_detect, dim3(bx, n), dim3(256), 0, s, desc + first, dtype, C,
                           W, ref_amp + (size_t)first * C * C, tgt_amp + (size_t)first * W * W);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_chip_stats(hipStream_t s, const float *amp, int pairs, int C, float *rT,
                             double2 *stats)
{
    hipLaunchKernelGGL(k_chip_stats, dim3(pairs), dim3(kStatThreads), 0, s, amp, C, rT, stats);
    return hipGetLastError();
}

hipError_t launch_sat(hipStream_t s, const float *amp, int pairs, int W, double *S1, double *S2)
{
    const dim3 grid((pairs + 3) / 4), block(256);
    switch ((W + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL(k_sat<1>, grid, block, 0, s, amp, W, pairs, S1, S2); break;
    case 2: hipLaunchKernelGGL(k_sat<2>, grid, block, 0, s, amp, W, pairs, S1, S2); break;
    case 3: hipLaunchKernelGGL(k_sat<3>, grid, block, 0, s, amp, W, pairs, S1, S2); break;
    case 4: hipLaunchKernelGGL(k_sat<4>, grid, block, 0, s, amp, W, pairs, S1, S2); break;
    case 5: hipLaunchKernelGGL(k_sat<5>, grid, block, 0, s, amp, W, pairs, S1, S2); break;
    case 6: hipLaunchKernelGGL(k_sat<6>, grid, block, 0, s, amp, W, pairs, S1, S2); break;
    case 7: hipLaunchKernelGGL(k_sat<7>, grid, block, 0, s, amp, W, pairs, S1, S2); break;
    case 8: hipLaunchKernelGGL(k_sat<8>, grid, block, 0, s, amp, W, pairs, S1, S2); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_correlate(hipStream_t s, const CorrelateShape &sh, const float *rT,
                            const float *amp, int pairs, int C, int W, float *partial)
{
    const int P = W - C + 1;
    const dim3 grid((unsigned)pairs * sh.tiles_u * sh.tiles_v * sh.nsplit), block(sh.threads);
#define AMPCOR_CORRELATE(BUV)                                                                    \
    hipLaunchKernelGGL(k_correlate<BUV>, grid, block, sh.lds_bytes, s, rT, amp, C, W, P, sh.ub,  \
                       sh.vb, sh.tiles_u, sh.tiles_v, sh.nsplit, partial)
    switch (sh.bu) {
    case 8: AMPCOR_CORRELATE(8); break;
    case 4: AMPCOR_CORRELATE(4); break;
    case 2: AMPCOR_CORRELATE(2); break;
    case 1: AMPCOR_CORRELATE(1); break;
    default: return hipErrorInvalidValue;
    }
#undef AMPCOR_CORRELATE
    return hipGetLastError();
}

hipError_t launch_normalize(hipStream_t s, const float *partial, int nsplit, const double *S1,
                            const double *S2, const double2 *stats, int pairs, int C, int W,
                            float *gamma)
{
    const int P = W - C + 1;
    const size_t total = (size_t)pairs * P * P;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(k_normalize, dim3(blocks), dim3(256), 0, s, partial, nsplit, S1, S2, stats,
                       C, W, P, pairs, gamma);
    return hipGetLastError();
}

hipError_t launch_coarse_peak(hipStream_t s, const float *gamma, int pairs, int P, int W, int E,
                              int refine_margin, int exclude, int2 *loc, int2 *window,
                              float *peak, float *snr)
{
    hipLaunchKernelGGL(k_coarse_peak, dim3((pairs + 3) / 4), dim3(256), 0, s, gamma, pairs, P, W,
                       E, refine_margin, exclude, loc, window, peak, snr);
    return hipGetLastError();
}

hipError_t launch_upsample(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int role,
                           const int2 *window, int N, int rf, const float *U, const float *Ut,
                           float2 *work, float *amp)
{
    const int M = rf * N;
    for (int first = 0; first < pairs; first += kMaxGridY) {
        const int n = pairs - first < kMaxGridY ? pairs - first : kMaxGridY;
        hipLaunchKernelGGL(k_upsample_rows, dim3((N + kUpRows - 1) / kUpRows, n), dim3(256),
                           kUpRows * N * sizeof(float2), s, desc + first, dtype, role,
                           window + first, N, rf, Ut, work + (size_t)first * N * M);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_upsample_cols, dim3((M + kUpCols - 1) / kUpCols, n), dim3(256),
                           N * kUpCols * sizeof(float2), s, work + (size_t)first * N * M, N, rf,
                           U, amp + (size_t)first * M * M);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_zoom(hipStream_t s, const float *gamma_refined, int pairs, int R, int zf,
                       const double *dirichlet, const int2 *window, const float *coarse_peak,
                       int margin, int rf, float *offsets, float *refined_peak, int2 *zoom_loc)
{
    const int span = 2 * zf + 1;
    const size_t lds = sizeof(double) * ((size_t)R * R + (size_t)span * R + (size_t)span * span);
    hipLaunchKernelGGL(k_zoom, dim3(pairs), dim3(kZoomThreads), lds, s, gamma_refined, R, zf,
                       dirichlet, window, coarse_peak, margin, rf, offsets, refined_peak,
                       zoom_loc);
    return hipGetLastError();
}

}  // namespace ampcor_hip
