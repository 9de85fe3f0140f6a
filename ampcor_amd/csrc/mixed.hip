// mixed.hip -- the mixed-precision form of the contraction on the matrix cores.
//
// The chip (zero-mean) and the target window (minus its mean) are scaled by
// powers of two to unit max-abs, rounded to BF16 or FP16 and correlated with
// v_mfma_f32_16x16x16 (FP32 accumulate).  The window statistics stay FP64
// (the same window sums as the FP32 form), and the normalizer puts the
// removed means back:
//   N(u, v) = sum r_q (t - mu_uv) = num / (a b) + (mu_w - mu_uv) sum r_q,
//   gamma   = N / (|r_q| sqrt(sum t^2 - mu_uv sum t)).
//
// The contraction is cast as sum_i T_i R_i: for chip row i, T_i holds target
// rows u + i (M = placement row u, K = target column x) and R_i is the banded
// Toeplitz matrix R_i[x][v] = r(i, x - v).  Target fragments are aligned
// 8-byte LDS reads; each Toeplitz fragment is four consecutive chip values
// starting at a lane-dependent offset, assembled from two aligned 8-byte
// reads with byte-align and selects.
//
// Replaces the reference's per-placement launches of correlate.cu:38-156 with
// one launch over all pairs; no counterpart exists for the precision modes.
#include "kernels.h"

#include <math.h>

namespace ampcor_hip {

namespace {

constexpr int kWave = 64;
constexpr int kPrepThreads = 256;

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef short shortx4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t to_bf16(float x)
{
    const uint32_t b = __float_as_uint(x);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline uint16_t to_f16(float x)
{
    const _Float16 h = (_Float16)x;
    return __builtin_bit_cast(uint16_t, h);
}

__device__ inline float from_f16(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

__device__ inline double block_sum(double *p, double x)
{
    const int t = threadIdx.x;
    p[t] = x;
    __syncthreads();
    for (int s = kPrepThreads / 2; s > 0; s >>= 1) {
        if (t < s) p[t] = p[t] + p[t + s];
        __syncthreads();
    }
    const double r = p[0];
    __syncthreads();
    return r;
}

__device__ inline float block_max(float *p, float x)
{
    const int t = threadIdx.x;
    p[t] = x;
    __syncthreads();
    for (int s = kPrepThreads / 2; s > 0; s >>= 1) {
        if (t < s) p[t] = fmaxf(p[t], p[t + s]);
        __syncthreads();
    }
    const float r = p[0];
    __syncthreads();
    return r;
}

// power of two that brings max-abs into [0.5, 1)
__device__ inline double unit_scale(float maxabs)
{
    if (!(maxabs > 0.0f)) return 1.0;
    int e;
    frexp((double)maxabs, &e);
    return ldexp(1.0, -e);
}

// ---------------------------------------------------------------------------
// per pair: quantized chip rows (zero padding of 16 on the left, CS wide),
// quantized target window, and the constants of the normalizer
// ---------------------------------------------------------------------------
template <bool FP16>
__global__ __launch_bounds__(kPrepThreads) void k_prep_mixed(
    const float *__restrict__ ref_amp, const float *__restrict__ tgt_amp, int C, int W, int CS,
    uint16_t *__restrict__ chip_q, uint16_t *__restrict__ tgt_q, MixedStats *__restrict__ stats,
    uint16_t *__restrict__ chip_lo, uint16_t *__restrict__ tgt_lo)
{
    __shared__ double pd[kPrepThreads];
    __shared__ float pf[kPrepThreads];
    const int pair = blockIdx.x;
    const int t = threadIdx.x;
    const float *r = ref_amp + (size_t)pair * C * C;
    const float *g = tgt_amp + (size_t)pair * W * W;
    uint16_t *cq = chip_q + (size_t)pair * C * CS;
    uint16_t *tq = tgt_q + (size_t)pair * W * W;

    double acc = 0.0;
    for (int k = t; k < C * C; k += kPrepThreads) acc += r[k];
    const float mean_r = (float)(block_sum(pd, acc) / ((double)C * C));
    float mx = 0.0f;
    for (int k = t; k < C * C; k += kPrepThreads) mx = fmaxf(mx, fabsf(r[k] - mean_r));
    const double a = unit_scale(block_max(pf, mx));
    double s1 = 0.0, s2 = 0.0;
    for (int k = t; k < C * CS; k += kPrepThreads) {
        const int i = k / CS, j = k - i * CS - 16;
        uint16_t q = 0, ql = 0;
        if (j >= 0 && j < C) {
            const float x = (float)((double)(r[(size_t)i * C + j] - mean_r) * a);
            q = FP16 ? to_f16(x) : to_bf16(x);
            const float hq = FP16 ? from_f16(q) : from_bf16(q);
            double v = (double)hq;
            if (chip_lo) {
                ql = FP16 ? to_f16(x - hq) : to_bf16(x - hq);
                v += (double)(FP16 ? from_f16(ql) : from_bf16(ql));
            }
            v /= a;
            s1 += v;
            s2 += v * v;
        }
        cq[k] = q;
        if (chip_lo) chip_lo[(size_t)pair * C * CS + k] = ql;
    }
    const double sumr = block_sum(pd, s1);
    const double sumr2 = block_sum(pd, s2);

    acc = 0.0;
    for (int k = t; k < W * W; k += kPrepThreads) acc += g[k];
    const float mean_t = (float)(block_sum(pd, acc) / ((double)W * W));
    mx = 0.0f;
    for (int k = t; k < W * W; k += kPrepThreads) mx = fmaxf(mx, fabsf(g[k] - mean_t));
    const double b = unit_scale(block_max(pf, mx));
    for (int k = t; k < W * W; k += kPrepThreads) {
        const float x = (float)((double)(g[k] - mean_t) * b);
        const uint16_t q = FP16 ? to_f16(x) : to_bf16(x);
        tq[k] = q;
        if (tgt_lo) {
            const float hq = FP16 ? from_f16(q) : from_bf16(q);
            tgt_lo[(size_t)pair * W * W + k] = FP16 ? to_f16(x - hq) : to_bf16(x - hq);
        }
    }
    if (t == 0) {
        MixedStats st;
        st.sumr = sumr;
        st.sumr2 = sumr2;
        st.scale_r = a;
        st.scale_t = b;
        st.mean_t = (double)mean_t;
        stats[pair] = st;
    }
}

// ---------------------------------------------------------------------------
// the contraction on MFMA.  Workgroup = (pair, block of u-tiles, block of
// v-tiles, chip-row range); wave w owns u-tile w of the block and every
// v-tile of it.  LDS: the target rows/columns the block needs, then the
// chip rows of the range.
// ---------------------------------------------------------------------------
template <bool FP16>
__device__ inline floatx4 mfma16(uint2 a, uint2 b, floatx4 c)
{
    if constexpr (FP16)
        return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(halfx4, a),
                                                     __builtin_bit_cast(halfx4, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(shortx4, a),
                                                         __builtin_bit_cast(shortx4, b), c, 0, 0, 0);
}

constexpr int kMaxVT = 5;

template <int NK, bool FP16, bool SPLIT>
__global__ __launch_bounds__(512) void k_correlate_mfma(
    const uint16_t *__restrict__ chip_q, const uint16_t *__restrict__ tgt_q, int C, int W, int P,
    int CS, int wg_ut, int wg_vt, int tiles_u, int tiles_v, int nsplit,
    float *__restrict__ partial, const uint16_t *__restrict__ chip_lo,
    const uint16_t *__restrict__ tgt_lo)
{
    extern __shared__ uint2 lds_u2[];
    uint16_t *lds = reinterpret_cast<uint16_t *>(lds_u2);
    int bidx = blockIdx.x;
    const int part = bidx % nsplit;
    bidx /= nsplit;
    const int tv = bidx % tiles_v;
    bidx /= tiles_v;
    const int tu = bidx % tiles_u;
    const int pair = bidx / tiles_u;

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int ut0 = tu * wg_ut, vt0 = tv * wg_vt;
    const int ilen = C / nsplit, i0 = part * ilen;
    const int nrows = 16 * wg_ut + ilen - 1;
    const int ncols = 16 * (wg_vt + NK - 1);
    const int TS = ncols + 8;  // row stride (elements), keeps 8-byte alignment
    constexpr int planes = SPLIT ? 2 : 1;
    uint16_t *tile = lds;                                      // [plane][nrows][TS]
    uint16_t *chip = lds + (size_t)planes * nrows * TS;        // [plane][ilen][CS]

    // stage the target block (zero outside the window) and the chip rows
    const int r0 = 16 * ut0 + i0, c0 = 16 * vt0;
#pragma unroll
    for (int pl = 0; pl < planes; ++pl) {
        const uint16_t *src = (pl ? tgt_lo : tgt_q) + (size_t)pair * W * W;
        uint16_t *dstt = tile + (size_t)pl * nrows * TS;
        for (int k = threadIdx.x; k < nrows * (ncols / 4); k += blockDim.x) {
            const int rr = k / (ncols / 4), cc = 4 * (k - rr * (ncols / 4));
            const int gr = r0 + rr, gc = c0 + cc;
            uint2 v = make_uint2(0u, 0u);
            if (gr < W) {
                uint16_t e[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) e[q] = (gc + q < W) ? src[(size_t)gr * W + gc + q] : 0;
                v.x = e[0] | ((uint32_t)e[1] << 16);
                v.y = e[2] | ((uint32_t)e[3] << 16);
            }
            *reinterpret_cast<uint2 *>(dstt + (size_t)rr * TS + cc) = v;
        }
        const uint16_t *csrc = (pl ? chip_lo : chip_q) + ((size_t)pair * C + i0) * CS;
        uint16_t *dstc = chip + (size_t)pl * ilen * CS;
        for (int k = threadIdx.x; k < ilen * CS / 4; k += blockDim.x)
            reinterpret_cast<uint2 *>(dstc)[k] = reinterpret_cast<const uint2 *>(csrc)[k];
    }
    __syncthreads();

    const int ut = ut0 + wave;
    const int nvt = min(wg_vt, tiles_v * wg_vt - vt0) ;
    if (wave >= wg_ut || 16 * ut >= P) return;
    const int vcount = min(nvt, (P - 16 * vt0 + 15) / 16);

    // two accumulators per tile (even and odd K-steps): twice the independent
    // MFMA chains; summed at the end
    floatx4 acc[kMaxVT], acc2[kMaxVT];
#pragma unroll
    for (int v = 0; v < kMaxVT; ++v) acc[v] = acc2[v] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    // Toeplitz fragment geometry: element index j = 4 D - beta + e,
    // D = 4 kappa + lk + 4 - alpha, lm = 4 alpha + beta
    const int alpha = lm >> 2, beta = lm & 3;
    const bool hi_pair = beta >= 2, odd = beta & 1;
    auto build = [&](const uint16_t *crow, uint2 *B) {
#pragma unroll
        for (int kap = 0; kap < NK; ++kap) {
            const int D = 4 * kap + lk + 4 - alpha;
            const uint2 lo = *reinterpret_cast<const uint2 *>(crow + 4 * (D - 1));
            const uint2 hi = *reinterpret_cast<const uint2 *>(crow + 4 * D);
            const uint32_t s0 = hi_pair ? lo.x : lo.y;
            const uint32_t s1 = hi_pair ? lo.y : hi.x;
            const uint32_t s2 = hi_pair ? hi.x : hi.y;
            B[kap].x = odd ? __builtin_amdgcn_alignbyte(s1, s0, 2) : s1;
            B[kap].y = odd ? __builtin_amdgcn_alignbyte(s2, s1, 2) : s2;
        }
    };
    for (int ii = 0; ii < ilen; ++ii) {
        uint2 B[NK], BL[SPLIT ? NK : 1];
        build(chip + (size_t)ii * CS, B);
        if constexpr (SPLIT) build(chip + (size_t)(ilen + ii) * CS, BL);
        const size_t arow_off = (size_t)(16 * wave + lm + ii) * TS + 4 * lk;
        uint2 A[kMaxVT + NK - 1], AL[SPLIT ? kMaxVT + NK - 1 : 1];
#pragma unroll
        for (int xi = 0; xi < kMaxVT + NK - 1; ++xi)
            if (xi < vcount + NK - 1) {
                A[xi] = *reinterpret_cast<const uint2 *>(tile + arow_off + 16 * xi);
                if constexpr (SPLIT)
                    AL[xi] = *reinterpret_cast<const uint2 *>(tile + (size_t)nrows * TS + arow_off + 16 * xi);
            }
#pragma unroll
        for (int kap = 0; kap < NK; ++kap)
#pragma unroll
            for (int v = 0; v < kMaxVT; ++v)
                if (v < vcount) {
                    floatx4 &c = (kap & 1) ? acc2[v] : acc[v];
                    c = mfma16<FP16>(A[v + kap], B[kap], c);
                    if constexpr (SPLIT) {
                        c = mfma16<FP16>(A[v + kap], BL[kap], c);
                        c = mfma16<FP16>(AL[v + kap], B[kap], c);
                    }
                }
    }
    float *dst = partial + ((size_t)pair * nsplit + part) * P * P;
#pragma unroll
    for (int v = 0; v < kMaxVT; ++v) {
        if (v >= vcount) break;
        const int vv = 16 * (vt0 + v) + lm;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int uu = 16 * ut + 4 * lk + e;
            if (uu < P && vv < P) dst[(size_t)uu * P + vv] = acc[v][e] + acc2[v][e];
        }
    }
}

__global__ __launch_bounds__(256) void k_normalize_mixed(const float *__restrict__ partial,
                                                         int nsplit,
                                                         const double2 *__restrict__ sums,
                                                         const MixedStats *__restrict__ stats,
                                                         int C, int P, int pairs,
                                                         float *__restrict__ gamma)
{
    const size_t total = (size_t)pairs * P * P;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int pp = P * P;
    const int pair = (int)(idx / pp);
    const int c = (int)(idx - (size_t)pair * pp);
    float num = 0.0f;
    for (int k = 0; k < nsplit; ++k) num += partial[((size_t)pair * nsplit + k) * pp + c];
    const double2 ws = sums[idx];
    const MixedStats st = stats[pair];
    const double n = (double)C * C;
    const double mu = ws.x / n;
    const double var = ws.y - mu * ws.x;
    // the numerator does not see a constant added to the chip, the norm is centred
    const double N = (double)num / (st.scale_r * st.scale_t) + (st.mean_t - mu) * st.sumr;
    const double c2 = st.sumr2 - st.sumr * st.sumr / n;
    const double B = (var > 0.0 && c2 > 0.0) ? sqrt(c2) * sqrt(var) : 0.0;
    gamma[idx] = (B > 0.0) ? (float)(N / B) : __builtin_nanf("");
}

}  // namespace

MixedShape choose_mixed_shape(int C, int P, int nsplit, bool split)
{
    MixedShape s{};
    s.nk = (C + 15) / 16 + 1;
    s.cs = 16 * (s.nk + 1);
    s.split = split;
    const int tiles = (P + 15) / 16;
    s.wg_vt = tiles < kMaxVT ? tiles : kMaxVT;
    s.wg_ut = tiles < 8 ? tiles : 8;
    s.nsplit = 1;
    const size_t budget = 120 * 1024;
    auto lds = [&](int ut, int vt, int ns) {
        const int ilen = C / ns;
        const size_t rows = 16 * ut + ilen - 1, cols = 16 * (vt + s.nk - 1) + 8;
        return (rows * cols + (size_t)ilen * s.cs) * sizeof(uint16_t) * (split ? 2 : 1);
    };
    while (lds(s.wg_ut, s.wg_vt, s.nsplit) > budget) {
        if (C % (2 * s.nsplit) == 0 && C / (2 * s.nsplit) >= 16) s.nsplit *= 2;
        else if (s.wg_ut > 1) s.wg_ut = (s.wg_ut + 1) / 2;
        else if (s.wg_vt > 1) s.wg_vt = (s.wg_vt + 1) / 2;
        else break;
    }
    // more independent, smaller workgroups for small surfaces: at most ~40 KiB
    // of LDS each, so that several share a CU
    while (tiles <= 3 && C % (2 * s.nsplit) == 0 && C / (2 * s.nsplit) >= 32 &&
           lds(s.wg_ut, s.wg_vt, s.nsplit) > 40 * 1024)
        s.nsplit *= 2;
    // a requested split (tuning) that keeps at least 16 chip rows per range
    if (nsplit > s.nsplit && (nsplit & (nsplit - 1)) == 0 && C % nsplit == 0 && C / nsplit >= 16)
        s.nsplit = nsplit;
    s.tiles_u = (tiles + s.wg_ut - 1) / s.wg_ut;
    s.tiles_v = (tiles + s.wg_vt - 1) / s.wg_vt;
    s.lds_bytes = lds(s.wg_ut, s.wg_vt, s.nsplit);
    return s;
}

hipError_t launch_prep_mixed(hipStream_t s, const float *ref_amp, const float *tgt_amp, int pairs,
                             int C, int W, int CS, bool fp16, uint16_t *chip_q, uint16_t *tgt_q,
                             MixedStats *stats, uint16_t *chip_lo, uint16_t *tgt_lo)
{
    if (fp16)
        hipLaunchKernelGGL(k_prep_mixed<true>, dim3(pairs), dim3(kPrepThreads), 0, s, ref_amp,
                           tgt_amp, C, W, CS, chip_q, tgt_q, stats, chip_lo, tgt_lo);
    else
        hipLaunchKernelGGL(k_prep_mixed<false>, dim3(pairs), dim3(kPrepThreads), 0, s, ref_amp,
                           tgt_amp, C, W, CS, chip_q, tgt_q, stats, chip_lo, tgt_lo);
    return hipGetLastError();
}

hipError_t launch_correlate_mixed(hipStream_t s, const MixedShape &sh, const uint16_t *chip_q,
                                  const uint16_t *tgt_q, int pairs, int C, int W, bool fp16,
                                  float *partial, const uint16_t *chip_lo, const uint16_t *tgt_lo)
{
    const int P = W - C + 1;
    if (sh.split && (!chip_lo || !tgt_lo)) return hipErrorInvalidValue;
    const dim3 grid((unsigned)pairs * sh.tiles_u * sh.tiles_v * sh.nsplit), block(64 * sh.wg_ut);
#define AMPCOR_MIXED_ONE(NKV, F16, SPL)                                                             \
    hipLaunchKernelGGL((k_correlate_mfma<NKV, F16, SPL>), grid, block, sh.lds_bytes, s, chip_q,     \
                       tgt_q, C, W, P, sh.cs, sh.wg_ut, sh.wg_vt, sh.tiles_u, sh.tiles_v,            \
                       sh.nsplit, partial, chip_lo, tgt_lo)
#define AMPCOR_MIXED(NKV)                                                                           \
    do {                                                                                            \
        if (sh.split) {                                                                             \
            if (fp16) AMPCOR_MIXED_ONE(NKV, true, true);                                            \
            else AMPCOR_MIXED_ONE(NKV, false, true);                                                \
        } else {                                                                                    \
            if (fp16) AMPCOR_MIXED_ONE(NKV, true, false);                                           \
            else AMPCOR_MIXED_ONE(NKV, false, false);                                               \
        }                                                                                           \
    } while (0)
    switch (sh.nk) {
    case 2: AMPCOR_MIXED(2); break;
    case 3: AMPCOR_MIXED(3); break;
    case 4: AMPCOR_MIXED(4); break;
    case 5: AMPCOR_MIXED(5); break;
    case 6: AMPCOR_MIXED(6); break;
    case 7: AMPCOR_MIXED(7); break;
    case 8: AMPCOR_MIXED(8); break;
    case 9: AMPCOR_MIXED(9); break;
    case 17: AMPCOR_MIXED(17); break;
    default: return hipErrorInvalidValue;
    }
#undef AMPCOR_MIXED
#undef AMPCOR_MIXED_ONE
    return hipGetLastError();
}

hipError_t launch_normalize_mixed(hipStream_t s, const float *partial, int nsplit,
                                  const double2 *sums, const MixedStats *stats, int pairs, int C,
                                  int W, float *gamma)
{
    const int P = W - C + 1;
    const size_t total = (size_t)pairs * P * P;
    hipLaunchKernelGGL(k_normalize_mixed, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       partial, nsplit, sums, stats, C, P, pairs, gamma);
    return hipGetLastError();
}

}  // namespace ampcor_hip
