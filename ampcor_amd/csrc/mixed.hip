// mixed.hip -- the mixed-precision form of the contraction on the matrix cores.
//
// The chip (zero-mean) and the target window (minus its mean) are scaled by
// powers of two to unit max-abs, rounded to BF16 or FP16 and correlated with
// v_mfma_f32_16x16x32 (FP32 accumulate).  The window statistics stay FP64
// (the same window sums as the FP32 form), and the normalizer puts the
// removed means back:
//   N(u, v) = sum r_q (t - mu_uv) = num / (a b) + (mu_w - mu_uv) sum r_q,
//   gamma   = N / (|r_q| sqrt(sum t^2 - mu_uv sum t)).
//
// The contraction is cast as sum_i T_i R_i: for chip row i, T_i holds target
// rows u + i (M = placement row u, K = target column x) and R_i is the banded
// Toeplitz matrix R_i[x][v] = r(i, x - v).
//
// Replaces the reference's per-placement launches of correlate.cu:38-156 with
// one launch over all pairs; no counterpart exists for the precision modes.
#include "mixed_correlate.h"
#include "packed_correlate.h"

#include <math.h>

#include <algorithm>

namespace ampcor_hip {

namespace {

constexpr int kWave = 64;
constexpr int kPrepThreads = 256;


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

MixedShape choose_mixed_shape(int C, int P, int nsplit, bool split, int max_vt)
{
    MixedShape s{};
    s.nk = (C + 15 + 15) / 16;
    s.cs = 16 * s.nk + 16;
    s.split = split;
    const int tiles = (P + 15) / 16;
    // v-tiles per workgroup: at most kMaxVT (or max_vt), spread evenly over the blocks
    const int vmax = max_vt > 0 && max_vt < mixed_detail::kMaxVT ? max_vt : mixed_detail::kMaxVT;
    const int vblocks = (tiles + vmax - 1) / vmax;
    s.wg_vt = (tiles + vblocks - 1) / vblocks;
    // u-tiles (waves) per workgroup: at most max_waves, spread evenly over the blocks
    const int max_ut = mixed_detail::max_waves(s.nk);
    const int ublocks = (tiles + max_ut - 1) / max_ut;
    s.wg_ut = (tiles + ublocks - 1) / ublocks;
    s.nsplit = 1;
    const size_t budget = 120 * 1024;
    auto lds = [&](int ut, int ns) { return mixed_lds_bytes(C, s.nk, s.wg_vt, ut, ns, split); };
    // chip-row ranges stay even (the contraction takes two rows per step)
    auto can_split = [&](int ns, int min_rows) {
        return C % (2 * ns) == 0 && C / (2 * ns) >= min_rows && (C / (2 * ns)) % 2 == 0;
    };
    while (lds(s.wg_ut, s.nsplit) > budget) {
        if (can_split(s.nsplit, 16)) s.nsplit *= 2;
        else if (s.wg_ut > 1) s.wg_ut = (s.wg_ut + 1) / 2;
        else break;
    }
    // more independent, smaller workgroups for small surfaces: at most ~40 KiB
    // of LDS each, so that several share a CU
    while (tiles <= 3 && can_split(s.nsplit, 32) && lds(s.wg_ut, s.nsplit) > 40 * 1024)
        s.nsplit *= 2;
    // a requested split (tuning) that keeps at least 16 chip rows per range
    if (nsplit > s.nsplit && (nsplit & (nsplit - 1)) == 0 && C % nsplit == 0 && C / nsplit >= 16 &&
        (C / nsplit) % 2 == 0)
        s.nsplit = nsplit;
    s.tiles_u = (tiles + s.wg_ut - 1) / s.wg_ut;
    s.tiles_v = (tiles + s.wg_vt - 1) / s.wg_vt;
    s.lds_bytes = lds(s.wg_ut, s.nsplit);
    return s;
}

MixedShape choose_mixed_stream_shape(int C, int P, int max_vt)
{
    MixedShape s{};
    const int nk = (C + 15 + 15) / 16, tiles = (P + 15) / 16;
    // whole chunks of chip rows; window rows of whole 16-byte loads (W = P + C - 1)
    if (C % mixed_detail::kStreamRows || (P + C - 1) % 8) return s;
    const int vmax = max_vt > 0 && max_vt < mixed_detail::kMaxVT ? max_vt : mixed_detail::kMaxVT;
    const int vblocks = (tiles + vmax - 1) / vmax;
    const int wg_vt = (tiles + vblocks - 1) / vblocks;
    const int max_ut = mixed_detail::max_waves(nk);
    const int ublocks = (tiles + max_ut - 1) / max_ut;
    const int wg_ut = (tiles + ublocks - 1) / ublocks;
    // each thread brings at most kStreamItems loads of a chunk
    if (mixed_stream_items(nk, wg_vt) > mixed_detail::kStreamItems * 64 * wg_ut) return s;
    s.nk = nk;
    s.cs = 16 * nk + 16;
    s.split = false;
    s.wg_ut = wg_ut;
    s.wg_vt = wg_vt;
    s.tiles_u = (tiles + wg_ut - 1) / wg_ut;
    s.tiles_v = (tiles + wg_vt - 1) / wg_vt;
    s.nsplit = 1;
    s.lds_bytes = mixed_stream_lds_bytes(nk, wg_vt, wg_ut);
    s.stream = true;
    return s;
}

MixedShape choose_mixed_packed_shape(int C, int P, int max_vt, int wpe, bool split, int nsplit)
{
    MixedShape s{};
    const int nk = (C + 15 + 15) / 16;
    if (P < 33 || (P - 1) % 32 || C % 16 || !packed_nk_built(nk)) return s;
    if (nsplit < 1 || (nsplit & (nsplit - 1)) || C % nsplit) return s;
    const int nb = (P - 1) / 32, vtt = 2 * nb + 1;
    // the (v-tiles, chunk rows, waves per SIMD) combinations, built where
    // their registers fit
    static const int kinds[][3] = {{2, 4, 3}, {3, 4, 3}, {3, 4, 2}, {5, 4, 2}, {2, 8, 3},
                                   {3, 8, 3}, {3, 8, 2}, {5, 8, 3}, {5, 8, 2}};
    int best = -1, best_cost = 1 << 30;
    for (int k = 0; k < (int)(sizeof(kinds) / sizeof(kinds[0])); ++k) {
        const int vt = kinds[k][0], ch = kinds[k][1], w = kinds[k][2];
        if ((max_vt > 0 && vt > max_vt) || (wpe && w != wpe)) continue;
        if (!packed_kind_fits(nk, vt, w)) continue;
        const int nvg = (vtt + vt - 1) / vt, waves = nb * nvg;
        if (waves > std::min(packed_detail::kMaxWaves, 4 * w) || (C / nsplit) % ch) continue;
        // one-wave workgroups refill four chip rows at a time (a smaller LDS
        // footprint, more workgroups per CU); larger ones eight
        if (ch != (waves == 1 ? 4 : 8)) continue;
        if (!packed_refill_fits(nk, vt, ch, nb, nvg)) continue;
        // fewest wasted v-tiles, then the higher occupancy, then fewer waves
        const int cost = 1000 * (nvg * vt - vtt) + 100 * (3 - w) + waves;
        if (cost < best_cost) best_cost = cost, best = k;
    }
    if (best < 0) return s;
    const int vt = kinds[best][0], nvg = (vtt + vt - 1) / vt;
    s.nk = nk;
    s.cs = 16 * nk + 16;
    s.split = split;
    s.wg_ut = nb * nvg;
    s.wg_vt = vt;
    s.tiles_u = nb;
    s.tiles_v = nvg;
    s.nsplit = nsplit;
    s.packed = kinds[best][1];
    s.wpe = kinds[best][2];
    s.lds_bytes = packed_lds_bytes(nk, vt, s.packed, nb, nvg);
    s.stream = false;
    return s;
}

MixedShape packed_fill_ranges(const MixedShape &sh, int C, int pairs, int maxsplit)
{
    // four waves per SIMD of the 256 CUs: below that, a launch of one- and
    // two-wave workgroups leaves SIMDs idle (measured on the scene's refined
    // contraction: 512 pairs 0.099 -> 0.050 ms in four ranges, 2048 pairs
    // 0.154 -> 0.124 ms in two; no change from 3072 pairs up)
    constexpr long kFillWaves = 4 * 4 * 256;
    MixedShape r = sh;
    if (!sh.packed || sh.nsplit < 1) return r;
    while ((long)pairs * r.wg_ut * r.nsplit < kFillWaves && 2 * r.nsplit <= maxsplit &&
           C % (2 * r.nsplit) == 0 && (C / (2 * r.nsplit)) % r.packed == 0)
        r.nsplit *= 2;
    return r;
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
    if (sh.split && (!chip_lo || !tgt_lo)) return hipErrorInvalidValue;
    if (sh.nk != (C + 15 + 15) / 16 || sh.cs != 16 * sh.nk + 16 || (C / sh.nsplit) % 2 != 0)
        return hipErrorInvalidValue;
    if (sh.stream &&
        (sh.split || sh.nsplit != 1 || C % mixed_detail::kStreamRows || W % 8 ||
         mixed_stream_items(sh.nk, sh.wg_vt) > mixed_detail::kStreamItems * 64 * sh.wg_ut ||
         sh.lds_bytes != mixed_stream_lds_bytes(sh.nk, sh.wg_vt, sh.wg_ut)))
        return hipErrorInvalidValue;
    if (sh.packed) {
        const int P = W - C + 1;
        if (sh.stream || sh.nsplit < 1 || C % sh.nsplit || P < 33 || (P - 1) % 32 || C % 16 ||
            (C / sh.nsplit) % sh.packed || sh.tiles_u != (P - 1) / 32 ||
            sh.wg_ut != sh.tiles_u * sh.tiles_v ||
            (sh.wpe != 2 && sh.wpe != 3) || !packed_nk_built(sh.nk) ||
            !packed_kind_fits(sh.nk, sh.wg_vt, sh.wpe) ||
            sh.wg_ut > std::min(packed_detail::kMaxWaves, 4 * sh.wpe) ||
            sh.tiles_v * sh.wg_vt < (P + 15) / 16 ||
            !packed_refill_fits(sh.nk, sh.wg_vt, sh.packed, sh.tiles_u, sh.tiles_v) ||
            sh.lds_bytes != packed_lds_bytes(sh.nk, sh.wg_vt, sh.packed, sh.tiles_u, sh.tiles_v))
            return hipErrorInvalidValue;
#define AMPCOR_PK(N)                                                                               \
    case N:                                                                                        \
        return launch_correlate_packed_nk<N>(s, sh, chip_q, tgt_q, pairs, C, W, fp16, partial,     \
                                             chip_lo, tgt_lo)
        switch (sh.nk) {
            AMPCOR_PK(2); AMPCOR_PK(3); AMPCOR_PK(4); AMPCOR_PK(5); AMPCOR_PK(6); AMPCOR_PK(7);
            AMPCOR_PK(8); AMPCOR_PK(9); AMPCOR_PK(10); AMPCOR_PK(11); AMPCOR_PK(17);
        default: return hipErrorInvalidValue;
        }
#undef AMPCOR_PK
    }
#define AMPCOR_NK(N)                                                                               \
    case N:                                                                                        \
        return launch_correlate_mixed_nk<N>(s, sh, chip_q, tgt_q, pairs, C, W, fp16, partial,      \
                                            chip_lo, tgt_lo)
    switch (sh.nk) {
        AMPCOR_NK(2); AMPCOR_NK(3); AMPCOR_NK(4); AMPCOR_NK(5);
        AMPCOR_NK(6); AMPCOR_NK(7); AMPCOR_NK(8); AMPCOR_NK(9);
        AMPCOR_NK(10); AMPCOR_NK(11); AMPCOR_NK(12); AMPCOR_NK(13);
        AMPCOR_NK(14); AMPCOR_NK(15); AMPCOR_NK(16); AMPCOR_NK(17);
    default: return hipErrorInvalidValue;
    }
#undef AMPCOR_NK
}

hipError_t launch_normalize_mixed(hipStream_t s, const float *partial, int nsplit,
                                  const double2 *sums, const MixedStats *stats, int pairs, int C,
                                  int W, float *gamma)
{
    const int P = W - C + 1;
    const size_t pp = (size_t)P * P;
    const int chunk = std::min(pairs, launch_pair_cap(pp));
    for (int first = 0; first < pairs; first += chunk) {
        const int n = std::min(chunk, pairs - first);
        hipLaunchKernelGGL(k_normalize_mixed, dim3((unsigned)((n * pp + 255) / 256)), dim3(256), 0,
                           s, partial + (size_t)first * nsplit * pp, nsplit, sums + first * pp,
                           stats + first, C, P, n, gamma + first * pp);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace ampcor_hip
