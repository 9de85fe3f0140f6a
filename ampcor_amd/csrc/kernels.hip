// kernels.hip -- gfx950 kernels of the batched tile correlator.
//
// The coarse stage (gather/detect, chip statistics, window sums, contraction,
// normalization, peak) rounds exactly where the CPU oracle's FP32 twin does, so
// its correlation surfaces are bit-identical: FP64 sums run in fixed orders
// (a 256-way strided pass plus a halving tree for chips, sliding column boxes
// and a chunked Kogge-Stone row scan for windows), the contraction is a fused
// multiply-add chain in
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

#include <algorithm>
#include <type_traits>

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

// two values as OCP e4m3 bytes (round to nearest even; clamped to the finite
// range first, where the conversion would give NaN), in the low 16 bits
__device__ inline uint32_t pack_fp8(float lo, float hi)
{
    lo = fminf(fmaxf(lo, -448.0f), 448.0f);
    hi = fminf(fmaxf(hi, -448.0f), 448.0f);
    return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(lo, hi, 0, false) & 0xffffu;
}

__device__ inline float dequant_fp8(uint32_t byte) { return __builtin_amdgcn_cvt_f32_fp8((int)byte, 0); }

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
// chip transposed (column-major), each chip column contiguous for the contraction
// ---------------------------------------------------------------------------
constexpr int kStatThreads = 256;

// the sum of one value per thread (NT threads) by a fixed halving tree
template <int NT>
__device__ inline double tree_sum(double *p, double x)
{
    const int t = threadIdx.x;
    p[t] = x;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) p[t] = p[t] + p[t + s];
        __syncthreads();
    }
    const double r = p[0];
    __syncthreads();
    return r;
}

// STAGE: the chip is read from global memory once, into LDS (C^2 floats), and
// the three passes read it there -- the same values in the same order, so the
// results are those of the unstaged kernel, bit for bit.  NT threads: the
// order of the sums is NT-way, so only NT = kStatThreads reproduces the FP32
// twin's coarse stage; the refined stage (checked by tolerance) may take 1024.
template <bool STAGE, int NT>
__global__ __launch_bounds__(NT) void k_chip_stats(const float *__restrict__ amp, int C,
                                                   float *__restrict__ rT,
                                                   double2 *__restrict__ stats)
{
    __shared__ double p1[NT];
    __shared__ float tile[32][33];
    extern __shared__ float staged[];
    const int pair = blockIdx.x;
    const int t = threadIdx.x;
    const int cells = C * C;
    const float *a = amp + (size_t)pair * cells;
    float *out = rT + (size_t)pair * cells;
    if constexpr (STAGE) {
        const float4 *a4 = reinterpret_cast<const float4 *>(a);
        float4 *s4 = reinterpret_cast<float4 *>(staged);
        for (int k = t; k < cells / 4; k += NT) s4[k] = a4[k];
        for (int k = cells / 4 * 4 + t; k < cells; k += NT) staged[k] = a[k];
        __syncthreads();
        a = staged;
    }

    double acc = 0.0;
    for (int k = t; k < cells; k += NT) acc = acc + (double)a[k];
    const double mean = tree_sum<NT>(p1, acc) / (double)cells;

    double a1 = 0.0, a2 = 0.0;
    for (int k = t; k < cells; k += NT) {
        const double dr = (double)(float)((double)a[k] - mean);
        a1 = a1 + dr;
        a2 = a2 + dr * dr;
    }
    const double s1 = tree_sum<NT>(p1, a1);
    const double s2 = tree_sum<NT>(p1, a2);
    if (t == 0) stats[pair] = make_double2(s1, sqrt(s2));

    // the zero-mean chip, transposed through LDS in 32 x 32 blocks
    const int c = t & 31, r0 = t >> 5;
    for (int bi = 0; bi < C; bi += 32)
        for (int bj = 0; bj < C; bj += 32) {
            for (int r = r0; r < 32; r += NT / 32) {
                const int i = bi + r, j = bj + c;
                tile[r][c] = (i < C && j < C) ? (float)((double)a[(size_t)i * C + j] - mean) : 0.0f;
            }
            __syncthreads();
            for (int r = r0; r < 32; r += NT / 32) {
                const int j = bj + r, i = bi + c;
                if (i < C && j < C) out[(size_t)j * C + i] = tile[c][r];
            }
            __syncthreads();
        }
}

// ---------------------------------------------------------------------------
// window sums of t and t^2 over every C x C placement, FP64, one wavefront per
// pair.  Each lane keeps the vertical box sums V(u, x) = sum_{i<C} t(u+i, x) of
// its columns and slides them down one row per placement row u
// (V <- (V + t(u+C-1, x)) - t(u-1, x)); each row of V then gets an inclusive
// Kogge-Stone scan across lanes (64-column chunks with a running carry), i.e.
// one row of a summed-area table, and the window sum is the difference of two
// of its entries.
// ---------------------------------------------------------------------------
constexpr int kMaxWindow = 512;

__device__ inline void wave_inclusive_scan(double &v1, double &v2, int lane)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const double o1 = __shfl_up(v1, d, kWave);
        const double o2 = __shfl_up(v2, d, kWave);
        if (lane >= d) {
            v1 = o1 + v1;
            v2 = o2 + v2;
        }
    }
}

template <int Q>
__global__ __launch_bounds__(kWave) void k_window_sums(const float *__restrict__ amp, int W, int C,
                                                       double2 *__restrict__ sums)
{
    __shared__ double sx1[kMaxWindow];
    __shared__ double sx2[kMaxWindow];
    const int lane = threadIdx.x;
    const int pair = blockIdx.x;
    const int P = W - C + 1;
    const float *t = amp + (size_t)pair * W * W;
    double2 *out = sums + (size_t)pair * P * P;

    double V1[Q], V2[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int x = lane + kWave * q;
        V1[q] = V2[q] = 0.0;
        if (x < W)
#pragma unroll 8
            for (int i = 0; i < C; ++i) {
                const double d = (double)t[(size_t)i * W + x];
                V1[q] = V1[q] + d;
                V2[q] = V2[q] + d * d;
            }
    }
    for (int u = 0; u < P; ++u) {
        if (u > 0) {
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int x = lane + kWave * q;
                if (x < W) {
                    const double a = (double)t[(size_t)(u + C - 1) * W + x];
                    const double b = (double)t[(size_t)(u - 1) * W + x];
                    V1[q] = (V1[q] + a) - b;
                    V2[q] = (V2[q] + a * a) - b * b;
                }
            }
        }
        double carry1 = 0.0, carry2 = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int x = lane + kWave * q;
            double v1 = (x < W) ? V1[q] : 0.0;
            double v2 = (x < W) ? V2[q] : 0.0;
            wave_inclusive_scan(v1, v2, lane);
            v1 = carry1 + v1;
            v2 = carry2 + v2;
            carry1 = __shfl(v1, kWave - 1, kWave);
            carry2 = __shfl(v2, kWave - 1, kWave);
            if (x < W) {
                sx1[x] = v1;
                sx2[x] = v2;
            }
        }
        __syncthreads();
        for (int v = lane; v < P; v += kWave) {
            const double s1 = sx1[v + C - 1] - (v > 0 ? sx1[v - 1] : 0.0);
            const double s2 = sx2[v + C - 1] - (v > 0 ? sx2[v - 1] : 0.0);
            out[(size_t)u * P + v] = make_double2(s1, s2);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// the contraction: num(u, v) = sum_j sum_i r'(i, j) t(u + i, v + j)
//
// One workgroup per (pair, placement tile, chip-column range), one task per
// thread: a lane owns BU consecutive placement rows at one placement column
// and walks the chip column by column (j outer, i inner); down a column it
// slides a BU-deep register window over the target rows, so every LDS read
// feeds BU fused multiply-adds.  The zero-mean chip is read with wave-uniform
// addresses, BU values at a time through vector buffer loads issued one block
// ahead (their counter is not the LDS reads').  The target streams through LDS
// in blocks of `jblock` chip columns: a block needs every target row of the
// tile but only vcount + jblock - 1 target columns, so the LDS footprint need
// not grow with the chip width (one block covers the whole range when it fits;
// the autotuner also weighs smaller blocks, which leave room for more
// workgroups per CU).  The block
// is stored column by column with an odd column stride: a lane's walk down a
// chip column reads consecutive words (one base address per column, the row
// in the instruction's offset), and the lanes of consecutive placement
// columns fall on distinct banks.  The accumulation order (j ascending,
// i ascending) is the same whatever the tiling.
//
// EXACT (the FP32 drop-in's coarse stage, one chip-column range): each chip
// column is its own FP32 fused multiply-add chain of C terms, started at
// zero, and the column sums are added in FP64 in column order.  A single FP32
// chain over all C^2 terms loses up to ~1e-5 of the background correlation
// on 128 x 128 chips, where the target amplitudes sit on a large mean.  The
// epilogue then normalizes in FP64 (as k_normalize) and writes gamma, so the
// FP64 numerator never passes through a float buffer.  Otherwise (split
// chip-column ranges) each lane writes its range's FP32 partial numerator.
// ---------------------------------------------------------------------------
constexpr int kJB = 16;  // column block when the whole chip range does not fit

// LDS column stride of the staged target block (odd; a stride of 2 mod 32
// measured slower: the larger block costs occupancy)
__host__ __device__ constexpr int correlate_stride(int rows)
{
    return rows | 1;
}
// dword 3 of a raw buffer resource on gfx9 targets (bounds-checked loads)
constexpr int kBufferWord3 = 0x00020000;

// BU consecutive floats at byte offset `off` of a buffer (wave-uniform offset)
template <int BU>
__device__ inline void chip_block(__amdgpu_buffer_rsrc_t rs, int off, float (&r)[BU])
{
    if constexpr (BU % 4 == 0) {
#pragma unroll
        for (int q = 0; q < BU / 4; ++q) {
            const float4 x = __builtin_bit_cast(
                float4, __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16 * q, 0, 0));
            r[4 * q] = x.x;
            r[4 * q + 1] = x.y;
            r[4 * q + 2] = x.z;
            r[4 * q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < BU; ++q)
            r[q] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off + 4 * q, 0, 0));
    }
}

// gamma from the FP64 numerator and the FP64 window sums / chip statistics:
// (num - mu_t sum r') / (|r'| sqrt(sum t^2 - mu_t sum t)); NaN where the
// window or the chip has no variance
__device__ inline float ncc_from_sums(double num, double2 ws, double2 st, double n)
{
    const double mu = ws.x / n;
    const double var = ws.y - mu * ws.x;
    const double A = mu * st.x;
    const double B = (var > 0.0 && st.y > 0.0) ? st.y * sqrt(var) : 0.0;
    return (B > 0.0) ? (float)((num - A) / B) : __builtin_nanf("");
}

template <int BU, bool EXACT>
__global__ __launch_bounds__(1024) void k_correlate(const float *__restrict__ rT,
                                                    const float *__restrict__ amp, int C, int W,
                                                    int P, int ub_extent, int vb_extent,
                                                    int tiles_u, int tiles_v, int nsplit, int jblock,
                                                    float *__restrict__ partial,
                                                    const double2 *__restrict__ sums,
                                                    const double2 *__restrict__ stats,
                                                    float *__restrict__ gamma)
{
    extern __shared__ float tile[];
    int b = xcd_block(blockIdx.x, gridDim.x);
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
    const int cols = vcount + jblock - 1;
    const int cs = correlate_stride(rows);

    const int task = threadIdx.x;
    const bool active = task < vcount * nub;
    const int v = active ? task % vcount : 0;
    const int ubk = active ? task / vcount : 0;
    const float *src = amp + (size_t)pair * W * W;
    const float *rp = rT + (size_t)pair * C * C;
    float acc[BU];
    double total[EXACT ? BU : 1];
#pragma unroll
    for (int a = 0; a < BU; ++a) acc[a] = 0.0f;
#pragma unroll
    for (int a = 0; a < (EXACT ? BU : 1); ++a) total[a] = 0.0;
    // the chip columns of this range are one run of values (column j at
    // rp + j C), taken in blocks of BU; a block is loaded (vector loads: the
    // LDS reads keep their own counter) while the one before it is used.  A
    // load past the pair's chip reads zeros (buffer bounds), so a block
    // fetched ahead needs no clamp; only blocks inside a column are used.
    const __amdgpu_buffer_rsrc_t crs =
        __builtin_amdgcn_make_buffer_rsrc((void *)rp, (short)0, C * C * 4, kBufferWord3);
    int off = j0 * C * 4;
    float ra[BU], rb[BU];
    chip_block<BU>(crs, off, ra);
    float w[BU];
    // chain steps i0 .. i0 + BU - 1 of the lane's column window tc
    auto step = [&](const float *tc, int i0, const float(&r)[BU]) {
#pragma unroll
        for (int s = 0; s < BU; ++s) {
            w[(s + BU - 1) % BU] = tc[i0 + s + BU - 1];
#pragma unroll
            for (int a = 0; a < BU; ++a) acc[a] = fmaf(r[s], w[(s + a) % BU], acc[a]);
        }
    };
    // the last n < BU steps of a column (C not a multiple of BU), their chip
    // values loaded one by one at off
    auto step_part = [&](const float *tc, int i0, int n) {
        float r[BU];
#pragma unroll
        for (int s = 0; s < BU; ++s)
            r[s] = s < n ? __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                         crs, off + 4 * s, 0, 0))
                         : 0.0f;
#pragma unroll
        for (int s = 0; s < BU; ++s) {
            if (s < n) {
                w[(s + BU - 1) % BU] = tc[i0 + s + BU - 1];
#pragma unroll
                for (int a = 0; a < BU; ++a) acc[a] = fmaf(r[s], w[(s + a) % BU], acc[a]);
            }
        }
    };

    // The LDS block is a ring of `cols` target columns: target column
    // v0 + j0 + g (g >= 0) sits in slot g mod cols.  The first pass stages
    // columns g < cols; each later pass (chip columns jb .. jb + jn - 1 of the
    // range) needs g in [jb, jb + vcount + jn - 1) and stages only the jn
    // columns entering, into the slots of the columns the previous pass was
    // the last to use.
    const int gbase = v0 + j0;
    for (int jb = 0; jb < jlen; jb += jblock) {
        const int jn = min(jblock, jlen - jb);
        __syncthreads();
        if (jb == 0) {
            // (r, c) of element k = r cols + c, advanced without a division
            const int dr = blockDim.x / cols, dc = blockDim.x - dr * cols;
            int r = threadIdx.x / cols, c = threadIdx.x - r * cols;
            for (; r < rows; r += dr, c += dc) {
                if (c >= cols) c -= cols, ++r;
                if (r >= rows) break;
                const int gr = u0 + r, gc = gbase + c;
                tile[c * cs + r] = (gr < W && gc < W) ? src[(size_t)gr * W + gc] : 0.0f;
            }
        } else {
            const int g0 = jb + vcount - 1;   // first entering column
            for (int k = threadIdx.x; k < rows * jn; k += blockDim.x) {
                const int r = k / jn, c = k - r * jn;
                const int gr = u0 + r, gc = gbase + g0 + c;
                tile[((g0 + c) % cols) * cs + r] =
                    (gr < W && gc < W) ? src[(size_t)gr * W + gc] : 0.0f;
            }
        }
        __syncthreads();
        if (!active) continue;
        int slot = (v + jb) % cols;
        for (int jj = 0; jj < jn; ++jj, slot = slot + 1 == cols ? 0 : slot + 1) {
            const float *tc = tile + (size_t)slot * cs + ubk * BU;
#pragma unroll
            for (int a = 0; a < BU - 1; ++a) w[a] = tc[a];
            int i0 = 0;
            for (; i0 + 2 * BU <= C; i0 += 2 * BU) {
                chip_block<BU>(crs, off + 4 * BU, rb);
                step(tc, i0, ra);
                chip_block<BU>(crs, off + 8 * BU, ra);
                step(tc, i0 + BU, rb);
                off += 8 * BU;
            }
            if (i0 + BU <= C) {  // one whole block left
                chip_block<BU>(crs, off + 4 * BU, rb);
                step(tc, i0, ra);
#pragma unroll
                for (int s = 0; s < BU; ++s) ra[s] = rb[s];
                off += 4 * BU;
                i0 += BU;
            }
            if (i0 < C) {  // part of a block ends the column
                const int n = C - i0;
                step_part(tc, i0, n);
                off += 4 * n;
                chip_block<BU>(crs, off, ra);   // the next column's first block
            }
            if constexpr (EXACT) {
#pragma unroll
                for (int a = 0; a < BU; ++a) {
                    total[a] = total[a] + (double)acc[a];
                    acc[a] = 0.0f;
                }
            }
        }
    }
    if (!active) return;
    if constexpr (EXACT) {
        const double2 st = stats[pair];
        const double n = (double)C * C;
        const size_t base = (size_t)pair * P * P;
#pragma unroll
        for (int a = 0; a < BU; ++a) {
            const int uu = ubk * BU + a;
            if (uu < ucount) {
                const size_t idx = base + (size_t)(u0 + uu) * P + v0 + v;
                gamma[idx] = ncc_from_sums(total[a], sums[idx], st, n);
            }
        }
    } else {
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
                                                   const double2 *__restrict__ sums,
                                                   const double2 *__restrict__ stats, int C,
                                                   int P, int pairs, float *__restrict__ gamma)
{
    const size_t total = (size_t)pairs * P * P;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int pp = P * P;
    const int pair = (int)(idx / pp);
    const int c = (int)(idx - (size_t)pair * pp);

    // fixed pairwise tree over the chip-column partials
    float part[16];
    for (int k = 0; k < nsplit; ++k) part[k] = partial[((size_t)pair * nsplit + k) * pp + c];
    for (int w = nsplit; w > 1; w >>= 1)
        for (int k = 0; k < w / 2; ++k) part[k] = part[2 * k] + part[2 * k + 1];
    const float num = part[0];

    gamma[idx] = ncc_from_sums((double)num, sums[idx], stats[pair], (double)C * C);
}

// ---------------------------------------------------------------------------
// window sums and normalization in one pass, for the stages that are checked
// against the FP64 oracle by tolerance (every mixed-precision stage and the
// FP32 refined stage; the FP32 coarse stage keeps the separate kernels whose
// order the CPU twin reproduces).  One workgroup per pair walks the
// placement rows in blocks of kFuseRows: threads own target columns and
// slide the vertical box sums V(u, x) = sum_{i<C} t(u+i, x) down in FP64;
// the block's rows of V go to LDS and become row prefix sums (16 threads a
// row, a 16-lane scan of the chunk totals), and each window sum is the
// difference of two prefix entries.  The numerator partials (chip-row or
// chip-column ranges) are summed in order, and the sums never leave the chip.
//
// The first box sums come from all threads (a chunk of columns and every
// G-th row each, summed in order through LDS); each block's target rows and
// numerator partials are fetched while the previous block is processed.
//
// Forms: kFuseFP32 (FP32 amplitudes; chip stats (sum r', |r'|)), kFuseMixed
// (FP32 amplitudes; the mixed normalizer with the operand scales), and
// kFuseBF16 / kFuseFP16 / kFuseFP8, which take the window sums from the
// quantized target itself.  With every sum in quantized units the scales and the
// removed constants cancel, and gamma is the normalized correlation of
// exactly the operands the contraction multiplied:
//   gamma = (num - sum R sum T / n) /
//           sqrt((sum R^2 - (sum R)^2 / n) (sum T^2 - (sum T)^2 / n)).
// ---------------------------------------------------------------------------
constexpr int kFuseRows = 16;
constexpr int kFuseThreads = 256;
constexpr int kFuseMaxCols = 2;  // target columns per thread: W <= 512
constexpr int kFuseOut = 9;      // outputs per thread and block: P <= 144
constexpr int kFuseMaxSplit = 4; // numerator partials loaded a block ahead
constexpr int kFuseScan = 6;     // LDS entries in flight per step of the row scans

constexpr int kFuseInit = 4;     // rows in flight per thread in the first box sums
constexpr size_t kLdsPerCu = 160 * 1024;   // LDS bytes of a gfx950 CU

// LDS row stride (doubles) of the normalizer's sums: at least W + 1 and 16 mod
// 32, so that two consecutive rows sit 32 banks apart and the half-wave
// accesses of two rows (16 lanes each, consecutive or 18- / 22-dword strided
// chunks) fall in distinct banks
__host__ __device__ inline int fuse_row_stride(int W) { return (W + 1 + 15) / 32 * 32 + 16; }

template <int FORM>
__device__ inline float fuse_load(const void *t, size_t k)
{
    if constexpr (FORM == kFuseBF16)
        return __uint_as_float((uint32_t) static_cast<const uint16_t *>(t)[k] << 16);
    else if constexpr (FORM == kFuseFP16)
        return (float)__builtin_bit_cast(_Float16, static_cast<const uint16_t *>(t)[k]);
    else if constexpr (FORM == kFuseFP8)
        return dequant_fp8(static_cast<const uint8_t *>(t)[k]);
    else
        return static_cast<const float *>(t)[k];
}

// bytes per target value, and values per chunk read at once (16 bytes; 8
// bytes for FP8, which keeps eight running sums per thread)
template <int FORM>
__host__ __device__ constexpr int fuse_elem_bytes()
{
    return FORM == kFuseFP8 ? 1 : (FORM == kFuseBF16 || FORM == kFuseFP16) ? 2 : 4;
}
template <int FORM>
__host__ __device__ constexpr int fuse_chunk()
{
    return fuse_elem_bytes<FORM>() == 4 ? 4 : 8;
}

// value q of a chunk of the target (w.z, w.w unused for FP8)
template <int FORM>
__device__ inline float fuse_unpack(const uint4 &w, int q)
{
    if constexpr (FORM == kFuseFP8) {
        const uint32_t d = q < 4 ? w.x : w.y;
        return dequant_fp8((d >> (8 * (q & 3))) & 0xffu);
    } else if constexpr (FORM == kFuseBF16 || FORM == kFuseFP16) {
        const uint32_t d2 = (q >> 1) < 2 ? ((q >> 1) ? w.y : w.x) : ((q >> 1) == 2 ? w.z : w.w);
        const uint32_t h = (d2 >> (16 * (q & 1))) & 0xffffu;
        return FORM == kFuseBF16 ? __uint_as_float(h << 16)
                                 : (float)__builtin_bit_cast(_Float16, (uint16_t)h);
    } else {
        const uint32_t d = q < 2 ? (q ? w.y : w.x) : (q == 2 ? w.z : w.w);
        return __uint_as_float(d);
    }
}

// 1 / sqrt(p) for p > 0 in FP64: the hardware estimate refined by two Newton
// steps, y <- y + y (1 - p y^2) / 2 (each at least doubles the correct bits;
// the result is rounded to FP32 by the callers)
__device__ inline double rsqrt_fp64(double p)
{
    double y = __builtin_amdgcn_rsq(p);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double e = fma(-p * y, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// LDS-only barrier: the waves of a workgroup exchange data through LDS alone
// (global memory is read-only input or write-once output here), so the
// global loads in flight -- the next block's rows -- need not complete
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0); vmcnt, expcnt untouched
    __builtin_amdgcn_s_barrier();
}

// LAZY: the outputs add each chunk's offset to the chunk prefixes they read,
// instead of a second pass over every entry (the same FP64 sums); it needs a
// few more registers, so its workgroups are held to three per CU -- the LDS of
// the wide windows it is chosen for allows no more anyway
template <int FORM, int COLS, int JO, bool LAZY>
__global__ __launch_bounds__(kFuseThreads, COLS == 1 && JO < 9 ? (LAZY ? 3 : 4) : 2) void k_normalize_fused(
    const void *__restrict__ target, int W, int C, const float *__restrict__ partial, int nsplit,
    const double2 *__restrict__ stats, const MixedStats *__restrict__ mstats,
    float *__restrict__ gamma, int xr)
{
    extern __shared__ double fl[];   // [2][kFuseRows + xr][ws]
    const int pair = blockIdx.x, t = threadIdx.x;
    const int P = W - C + 1, ws = fuse_row_stride(W);
    // the lone-row path is built into the lazy-offset kernels only: in the
    // others its extra scan and outputs cost registers (spills at the
    // four-workgroup 128-VGPR cap, a workgroup per CU in the wide ones)
    if constexpr (!LAZY) xr = 0;
    double *p1 = fl, *p2 = fl + (kFuseRows + xr) * ws;
    const size_t tbase = (size_t)pair * W * W;
    auto tval = [&](int row, int x) { return fuse_load<FORM>(target, tbase + (size_t)row * W + x); };
    const double n = (double)C * C, inv_n = 1.0 / n;
    // per-pair constants: the chip sum and the squared centred chip norm in the
    // units of the target sums, and for the FP32-amplitude mixed form the
    // operand scales and the window constant
    double sumr, chip2, inv_ab = 1.0, mean_t = 0.0;
    if constexpr (FORM == kFuseFP32) {
        const double2 st = stats[pair];
        sumr = st.x;
        chip2 = st.y * st.y;
    } else {
        const MixedStats st = mstats[pair];
        if constexpr (FORM == kFuseMixed) {
            sumr = st.sumr;
            chip2 = st.sumr2 - st.sumr * st.sumr / n;
            inv_ab = 1.0 / (st.scale_r * st.scale_t);
            mean_t = st.mean_t;
        } else {
            sumr = st.sumr * st.scale_r;
            chip2 = st.sumr2 * st.scale_r * st.scale_r - sumr * sumr / n;
        }
    }

    // the rows entering (u + C - 1) and leaving (u - 1) the box of each of a
    // block's placement rows u = u0 + k, for this thread's columns
    // (rows and columns clamped into the window: every load is in bounds and
    // unconditional; the values of rows past the last placement are unused)
    auto fetch = [&](int u0, float (&ra)[COLS][kFuseRows], float (&rb)[COLS][kFuseRows]) {
#pragma unroll
        for (int q = 0; q < COLS; ++q) {
            const int x = min(t + kFuseThreads * q, W - 1);
#pragma unroll
            for (int k = 0; k < kFuseRows; ++k) {
                const int u = u0 + k;
                ra[q][k] = tval(min(u + C - 1, W - 1), x);
                rb[q][k] = tval(max(min(u - 1, W - 1), 0), x);
            }
        }
    };
    float ra[COLS][kFuseRows], rb[COLS][kFuseRows];
    fetch(0, ra, rb);
    // vertical box sums at u = 0 (rows [0, C)): each thread takes a 16-byte
    // chunk of columns (CH values) and every G-th row of it, kFuseInit rows in
    // flight at a time; the G row groups are then summed in order through LDS
    constexpr int CH = fuse_chunk<FORM>();
    double v1[COLS], v2[COLS];
    {
        const int ncc = (W + CH - 1) / CH;
        const int G = min(kFuseRows, kFuseThreads / ncc);
        const int g = t / ncc, cc = t - g * ncc;
        if (g < G) {
            double s1[CH], s2[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q) s1[q] = s2[q] = 0.0;
            const bool vec = W % CH == 0;
            for (int i0 = g; i0 < C; i0 += kFuseInit * G) {
                uint4 e[kFuseInit];   // raw chunks: four registers each
#pragma unroll
                for (int b = 0; b < kFuseInit; ++b) {
                    const int i = min(i0 + b * G, C - 1);   // clamped: in bounds, unused
                    const size_t at = tbase + (size_t)i * W;
                    if (vec) {
                        const char *src = static_cast<const char *>(target) +
                                          (at + CH * cc) * fuse_elem_bytes<FORM>();
                        if constexpr (FORM == kFuseFP8) {
                            const uint2 w2 = *reinterpret_cast<const uint2 *>(src);
                            e[b] = make_uint4(w2.x, w2.y, 0u, 0u);
                        } else {
                            e[b] = *reinterpret_cast<const uint4 *>(src);
                        }
                    } else {
                        uint32_t v[CH];
#pragma unroll
                        for (int q = 0; q < CH; ++q) {
                            const size_t k = at + min(CH * cc + q, W - 1);
                            if constexpr (FORM == kFuseFP8)
                                v[q] = static_cast<const uint8_t *>(target)[k];
                            else if constexpr (FORM == kFuseBF16 || FORM == kFuseFP16)
                                v[q] = static_cast<const uint16_t *>(target)[k];
                            else
                                v[q] = __float_as_uint(static_cast<const float *>(target)[k]);
                        }
                        if constexpr (FORM == kFuseFP8)
                            e[b] = make_uint4(v[0] | v[1] << 8 | v[2] << 16 | v[3] << 24,
                                              v[4] | v[5] << 8 | v[6] << 16 | v[7] << 24, 0u, 0u);
                        else if constexpr (CH == 8)
                            e[b] = make_uint4(v[0] | v[1] << 16, v[2] | v[3] << 16,
                                              v[4] | v[5] << 16, v[6] | v[7] << 16);
                        else
                            e[b] = make_uint4(v[0], v[1], v[2], v[3]);
                    }
                }
#pragma unroll
                for (int b = 0; b < kFuseInit; ++b)
                    if (i0 + b * G < C)
#pragma unroll
                        for (int q = 0; q < CH; ++q) {
                            const double d = (double)fuse_unpack<FORM>(e[b], q);
                            s1[q] = s1[q] + d;
                            s2[q] = s2[q] + d * d;
                        }
            }
#pragma unroll
            for (int q = 0; q < CH; ++q)
                if (CH * cc + q < W) {
                    p1[g * ws + CH * cc + q] = s1[q];
                    p2[g * ws + CH * cc + q] = s2[q];
                }
        }
        lds_barrier();
#pragma unroll
        for (int q = 0; q < COLS; ++q) {
            const int x = t + kFuseThreads * q;
            v1[q] = v2[q] = 0.0;
            if (x < W)
                for (int h = 0; h < G; ++h) {
                    v1[q] = v1[q] + p1[h * ws + x];
                    v2[q] = v2[q] + p2[h * ws + x];
                }
        }
        lds_barrier();   // the first block rewrites p1, p2
    }
    const int r = t >> 4, sl = t & 15;      // row of the block, slot in the row
    // the partials of a block's numerators, loaded a block ahead when there are
    // at most kFuseMaxSplit of them (row and column clamped: in bounds, unused
    // past the last placement)
    const bool ahead = nsplit <= kFuseMaxSplit;
    float raw[kFuseMaxSplit][JO];
    auto fetch_num = [&](int u0) {
        const float *pp = partial + (size_t)pair * nsplit * P * P + (size_t)min(u0 + r, P - 1) * P;
#pragma unroll
        for (int k = 0; k < kFuseMaxSplit; ++k)
#pragma unroll
            for (int j = 0; j < JO; ++j)
                raw[k][j] = k < nsplit ? pp[(size_t)k * P * P + min(sl + 16 * j, P - 1)] : 0.0f;
    };
    if (ahead) fetch_num(0);
    // prefix chunk per slot, odd: the 16 slots of a row start in distinct LDS banks
    const int cw = ((W + 15) / 16) | 1;
    const int x0 = min(W, sl * cw), x1 = min(W, x0 + cw);
    // (lazy form) chunk of column x = (x * inv_cw) >> 20, exact for x < 512 and cw <= 33
    const int inv_cw = (1 << 20) / cw + 1;
    __shared__ double off1[LAZY ? kFuseRows + 1 : 1][16], off2[LAZY ? kFuseRows + 1 : 1][16];

    for (int u0 = 0; u0 < P; u0 += kFuseRows) {
        const int rows = min(kFuseRows, P - u0);
        // xr: a lone last placement row joins the block before it as its row
        // kFuseRows (the LDS holds one row more); the threads of row 0 take
        // its scan and outputs after their own, which saves a whole block's
        // passes and barriers (every P = 16 n + 1)
        const bool lone = LAZY && xr && P - u0 == kFuseRows + 1;
        // the numerators of this thread's outputs (row r, columns sl + 16 j):
        // the partials fetched a block ahead, summed in split order, then the
        // next block's fetched
        float num[JO];
        if (ahead) {
#pragma unroll
            for (int j = 0; j < JO; ++j) {
                num[j] = raw[0][j];
#pragma unroll
                for (int k = 1; k < kFuseMaxSplit; ++k)
                    if (k < nsplit) num[j] += raw[k][j];
            }
            fetch_num(u0 + kFuseRows);
        } else {
#pragma unroll
            for (int j = 0; j < JO; ++j) num[j] = 0.0f;
            if (r < rows) {
                const float *pp = partial + (size_t)pair * nsplit * P * P + (size_t)(u0 + r) * P + sl;
                for (int k = 0; k < nsplit; ++k)
#pragma unroll
                    for (int j = 0; j < JO; ++j)
                        if (sl + 16 * j < P) num[j] += pp[(size_t)k * P * P + 16 * j];
            }
        }
        // slide V down to the block's rows and park them in LDS
#pragma unroll
        for (int q = 0; q < COLS; ++q) {
            const int x = t + kFuseThreads * q;
            if (x >= W) continue;
#pragma unroll
            for (int k = 0; k < kFuseRows; ++k) {
                if (k >= rows) break;
                if (u0 + k > 0) {
                    const double da = ra[q][k], db = rb[q][k];
                    v1[q] = (v1[q] + da) - db;
                    v2[q] = (v2[q] + da * da) - db * db;
                }
                p1[k * ws + x] = v1[q];
                p2[k * ws + x] = v2[q];
            }
        }
        // the next block's rows, in flight during this block's scans and outputs
        fetch(u0 + kFuseRows, ra, rb);
        if (lone) {
            // the lone row u0 + kFuseRows: its entering and leaving rows are
            // the first ones just fetched (the slide of the block it would be)
#pragma unroll
            for (int q = 0; q < COLS; ++q) {
                const int x = t + kFuseThreads * q;
                if (x >= W) continue;
                const double da = ra[q][0], db = rb[q][0];
                v1[q] = (v1[q] + da) - db;
                v2[q] = (v2[q] + da * da) - db * db;
                p1[kFuseRows * ws + x] = v1[q];
                p2[kFuseRows * ws + x] = v2[q];
            }
        }
        lds_barrier();
        // row prefix sums of row rr: chunk prefix, then the 16-lane scan of
        // chunk totals (kFuseScan entries in flight per step); every lane of
        // the 16-lane group takes part in the scan, live or not
        auto scan_row = [&](int rr, bool live) {
            double run1 = 0.0, run2 = 0.0;
            if (live)
                for (int xb = x0; xb < x1; xb += kFuseScan) {
                    double e1[kFuseScan], e2[kFuseScan];
#pragma unroll
                    for (int i = 0; i < kFuseScan; ++i) {
                        e1[i] = xb + i < x1 ? p1[rr * ws + xb + i] : 0.0;
                        e2[i] = xb + i < x1 ? p2[rr * ws + xb + i] : 0.0;
                    }
#pragma unroll
                    for (int i = 0; i < kFuseScan; ++i) {
                        run1 += e1[i];
                        run2 += e2[i];
                        if (xb + i < x1) {
                            p1[rr * ws + xb + i] = run1;
                            p2[rr * ws + xb + i] = run2;
                        }
                    }
                }
            double o1 = run1, o2 = run2;
#pragma unroll
            for (int d = 1; d < 16; d <<= 1) {
                const double y1 = __shfl_up(o1, d, 16), y2 = __shfl_up(o2, d, 16);
                if (sl >= d) {
                    o1 += y1;
                    o2 += y2;
                }
            }
            o1 -= run1;   // exclusive: the totals of the chunks to the left
            o2 -= run2;
            if (LAZY && live) {
                off1[rr][sl] = o1;
                off2[rr][sl] = o2;
            }
            if (!LAZY && live && sl > 0)
                for (int xb = x0; xb < x1; xb += kFuseScan) {
                    double e1[kFuseScan], e2[kFuseScan];
#pragma unroll
                    for (int i = 0; i < kFuseScan; ++i) {
                        e1[i] = xb + i < x1 ? p1[rr * ws + xb + i] : 0.0;
                        e2[i] = xb + i < x1 ? p2[rr * ws + xb + i] : 0.0;
                    }
#pragma unroll
                    for (int i = 0; i < kFuseScan; ++i)
                        if (xb + i < x1) {
                            p1[rr * ws + xb + i] = e1[i] + o1;
                            p2[rr * ws + xb + i] = e2[i] + o2;
                        }
                }
        };
        scan_row(r, r < rows);
        if (lone && r == 0) scan_row(kFuseRows, true);
        lds_barrier();
        // the outputs of row rr of the block from its numerators nm
        auto emit_row = [&](int rr, const float(&nm)[JO]) {
            const int r = rr;   // the LDS row
            const float *num = nm;
            const int u = u0 + r;
            for (int j = 0; j < JO; ++j) {
                const int v = sl + 16 * j;
                if (v >= P) break;
                double S1, S2;
                if constexpr (LAZY) {
                    // the row's prefix sums: chunk prefix + chunk offset
                    const int xa = v + C - 1, ca = (xa * inv_cw) >> 20;
                    S1 = p1[r * ws + xa] + off1[r][ca];
                    S2 = p2[r * ws + xa] + off2[r][ca];
                    if (v > 0) {
                        const int cb = ((v - 1) * inv_cw) >> 20;
                        S1 = S1 - (p1[r * ws + v - 1] + off1[r][cb]);
                        S2 = S2 - (p2[r * ws + v - 1] + off2[r][cb]);
                    } else {
                        S1 = S1 - 0.0;
                        S2 = S2 - 0.0;
                    }
                } else {
                    S1 = p1[r * ws + v + C - 1] - (v > 0 ? p1[r * ws + v - 1] : 0.0);
                    S2 = p2[r * ws + v + C - 1] - (v > 0 ? p2[r * ws + v - 1] : 0.0);
                }
                // FP64 throughout: the centred sums cancel, and a tiny variance
                // must not flush to zero
                // (the BF16 / FP16 / mixed forms multiply by 1 / n: one FP64
                // rounding more, far below their operand rounding)
                const double mu = FORM == kFuseFP32 ? S1 / n : S1 * inv_n;
                const double var = S2 - mu * S1;
                const double num_c = FORM == kFuseMixed
                                         ? (double)num[j] * inv_ab + (mean_t - mu) * sumr
                                         : (double)num[j] - mu * sumr;
                float g = __builtin_nanf("");
                if (var > 0.0 && chip2 > 0.0) g = (float)(num_c * rsqrt_fp64(chip2 * var));
                gamma[(size_t)pair * P * P + (size_t)u * P + v] = g;
            }
        };
        if (r < rows) emit_row(r, num);
        if (lone && r == 0) {
            // the lone row's numerators, summed in split order: fetched a block
            // ahead (row 0's next block is the lone row), or read now
            float nl[JO];
            if (ahead) {
#pragma unroll
                for (int j = 0; j < JO; ++j) {
                    nl[j] = raw[0][j];
#pragma unroll
                    for (int k = 1; k < kFuseMaxSplit; ++k)
                        if (k < nsplit) nl[j] += raw[k][j];
                }
            } else {
                const float *pp =
                    partial + (size_t)pair * nsplit * P * P + (size_t)(u0 + kFuseRows) * P + sl;
#pragma unroll
                for (int j = 0; j < JO; ++j) nl[j] = 0.0f;
                for (int k = 0; k < nsplit; ++k)
#pragma unroll
                    for (int j = 0; j < JO; ++j)
                        if (sl + 16 * j < P) nl[j] += pp[(size_t)k * P * P + 16 * j];
            }
            emit_row(kFuseRows, nl);
        }
        lds_barrier();
        if (lone) break;
    }
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

// the wave's SNR sums reduced, then lane 0 writes the peak, SNR and nudged window
__device__ inline void coarse_peak_finish(int pair, int lane, int P, int W, int E, int rm, bool none,
                                          int row, int col, float pk, double sum, int count,
                                          int ties, int2 *__restrict__ loc,
                                          int2 *__restrict__ window, float *__restrict__ peak,
                                          float *__restrict__ snr, int *__restrict__ near)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        sum += __shfl_xor(sum, d, kWave);
        count += __shfl_xor(count, d, kWave);
        ties += __shfl_xor(ties, d, kWave);
    }
    if (lane == 0) {
        if (near) near[pair] = none ? 0 : ties;
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

// The surface has at most KEEP cells per lane and stays in registers between
// the argmax and the SNR sums (one read of it); each lane visits its cells in
// increasing order, so its ties resolve to the first.  near (optional): the
// number of cells within tau of the peak (the peak included), for the FP8
// peak guard.
template <int KEEP>
__global__ __launch_bounds__(256) void k_coarse_peak(const float *__restrict__ gamma, int pairs,
                                                     int P, int W, int E, int rm, int exclude,
                                                     int2 *__restrict__ loc,
                                                     int2 *__restrict__ window,
                                                     float *__restrict__ peak,
                                                     float *__restrict__ snr, float tau,
                                                     int *__restrict__ near)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int pair = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pair >= pairs) return;
    const int cells = P * P;
    const float *g = gamma + (size_t)pair * cells;
    static_assert(KEEP > 0, "surfaces of more cells take k_coarse_peak_wide");
    {
        // every load in flight at once; NaN past the last cell never wins and is
        // never summed
        float x[KEEP];
#pragma unroll
        for (int j = 0; j < KEEP; ++j) {
            const int k = lane + j * kWave;
            x[j] = k < cells ? g[k] : __builtin_nanf("");
        }
        float best = -INFINITY;
        int bidx = 0x7fffffff;
#pragma unroll
        for (int j = 0; j < KEEP; ++j)
            if (x[j] > best) {
                best = x[j];
                bidx = lane + j * kWave;
            }
        wave_argmax(best, bidx);
        const bool none = (bidx == 0x7fffffff);
        const int row = none ? 0 : bidx / P;
        const int col = none ? 0 : bidx - row * P;
        double sum = 0.0;
        int count = 0, ties = 0;
        const float floor_ = best - tau;
        // (u, v) of cell lane + 64 j, stepped by 64 cells without divisions
        const int du = kWave / P, dv = kWave - du * P;
        int u = lane / P, v = lane - u * P;
#pragma unroll
        for (int j = 0; j < KEEP; ++j) {
            ties += x[j] >= floor_;
            if (x[j] == x[j] && !(abs(u - row) <= exclude && abs(v - col) <= exclude)) {
                sum += (double)x[j] * (double)x[j];
                ++count;
            }
            u += du;
            v += dv;
            if (v >= P) {
                v -= P;
                ++u;
            }
        }
        coarse_peak_finish(pair, lane, P, W, E, rm, none, row, col, none ? __builtin_nanf("") : best,
                           sum, count, ties, loc, window, peak, snr, near);
        return;
    }
}

// Surfaces above KEEP x 64 cells: one workgroup (four waves) per pair.  Each
// thread visits cells t, t + 256, ... in increasing order (its ties resolve
// to the first, and the cell coordinates step without divisions); the
// waves' maxima and SNR sums meet in LDS.  KEEP > 0: the thread's cells (at
// most KEEP) stay in registers between the two passes; KEEP = 0 reads them twice.
template <int KEEP>
__global__ __launch_bounds__(256) void k_coarse_peak_wide(const float *__restrict__ gamma, int P,
                                                          int W, int E, int rm, int exclude,
                                                          int2 *__restrict__ loc,
                                                          int2 *__restrict__ window,
                                                          float *__restrict__ peak,
                                                          float *__restrict__ snr, float tau,
                                                          int *__restrict__ near)
{
    constexpr int kT = 256, kB = 8;   // threads; loads in flight per thread
    __shared__ float s_best[kT / kWave];
    __shared__ int s_idx[kT / kWave], s_count[kT / kWave], s_ties[kT / kWave];
    __shared__ double s_sum[kT / kWave];
    const int t = threadIdx.x, lane = t & (kWave - 1), wv = t >> 6;
    const int pair = blockIdx.x;
    const int cells = P * P;
    const float *g = gamma + (size_t)pair * cells;
    float best = -INFINITY;
    int bidx = 0x7fffffff;
    // (KEEP > 0) every load in flight at once; NaN past the last cell never
    // wins and is never summed
    float xk[KEEP > 0 ? KEEP : 1];
    if constexpr (KEEP > 0) {
#pragma unroll
        for (int j = 0; j < KEEP; ++j) {
            const int k = t + j * kT;
            xk[j] = k < cells ? g[k] : __builtin_nanf("");
        }
#pragma unroll
        for (int j = 0; j < KEEP; ++j)
            if (xk[j] > best) {
                best = xk[j];
                bidx = t + j * kT;
            }
    } else {
        for (int k0 = t; k0 < cells; k0 += kB * kT) {
            float x[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const int k = k0 + j * kT;
                x[j] = k < cells ? g[k] : -INFINITY;
            }
#pragma unroll
            for (int j = 0; j < kB; ++j)
                if (x[j] > best) {
                    best = x[j];
                    bidx = k0 + j * kT;
                }
        }
    }
    wave_argmax(best, bidx);
    if (lane == 0) {
        s_best[wv] = best;
        s_idx[wv] = bidx;
    }
    __syncthreads();
    best = s_best[0];
    bidx = s_idx[0];
#pragma unroll
    for (int w = 1; w < kT / kWave; ++w) better(best, bidx, s_best[w], s_idx[w]);
    const bool none = (bidx == 0x7fffffff);
    const int row = none ? 0 : bidx / P;
    const int col = none ? 0 : bidx - row * P;

    double sum = 0.0;
    int count = 0, ties = 0;
    const float floor_ = best - tau;
    const int du = kT / P, dv = kT - du * P;
    int u = t / P, v = t - u * P;
    // cell x at (u, v); then (u, v) of the thread's next cell
    auto visit = [&](float x) {
        ties += x >= floor_;
        if (x == x && !(abs(u - row) <= exclude && abs(v - col) <= exclude)) {
            sum += (double)x * (double)x;
            ++count;
        }
        u += du;
        v += dv;
        if (v >= P) {
            v -= P;
            ++u;
        }
    };
    if constexpr (KEEP > 0) {
#pragma unroll
        for (int j = 0; j < KEEP; ++j) visit(xk[j]);
    } else {
        for (int k0 = t; k0 < cells; k0 += kB * kT) {
            float x[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const int k = k0 + j * kT;
                x[j] = k < cells ? g[k] : __builtin_nanf("");
            }
#pragma unroll
            for (int j = 0; j < kB; ++j) visit(x[j]);
        }
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        sum += __shfl_xor(sum, d, kWave);
        count += __shfl_xor(count, d, kWave);
        ties += __shfl_xor(ties, d, kWave);
    }
    if (lane == 0) {
        s_sum[wv] = sum;
        s_count[wv] = count;
        s_ties[wv] = ties;
    }
    __syncthreads();
    if (wv != 0) return;
    // wave 0 finishes: lanes 0 .. 3 hold the waves' sums, the others nothing
    const bool mine = lane < kT / kWave;
    coarse_peak_finish(pair, lane, P, W, E, rm, none, row, col, none ? __builtin_nanf("") : best,
                       mine ? s_sum[lane] : 0.0, mine ? s_count[lane] : 0,
                       mine ? s_ties[lane] : 0, loc, window, peak, snr, near);
}

// ---------------------------------------------------------------------------
// FP8 peak guard.  The e4m3 operands move each coarse gamma by up to ~0.2 / C
// (measured on the 128 x 128 amplitude workload: 1.7e-3), so where the top of
// the surface is flatter than that the FP8 argmax may not be the FP64 one.
// Every placement whose FP8 gamma is within tau of the FP8 peak is a
// candidate; with more than one, each is re-scored by the FP64 normalized
// correlation of the amplitudes, read straight from the rasters, and the best
// (ties: the first in row-major order) becomes the peak location.  Values stay
// those of the FP8 surface: the peak is the surface at the chosen cell and the
// SNR is recomputed around it.  One workgroup per pair.
// ---------------------------------------------------------------------------
constexpr int kGuardThreads = 256;
constexpr int kGuardMax = 16;  // candidates re-scored per pair

__device__ inline double detect64(int dtype, const void *base, size_t offset)
{
    if (dtype == 0) {
        const float2 z = static_cast<const float2 *>(base)[offset];
        const double re = z.x, im = z.y;
        return sqrt(re * re + im * im);
    }
    return fabs((double)static_cast<const float *>(base)[offset]);
}

// sum of x over the workgroup (fixed tree; every thread gets the result)
template <int N>
__device__ inline void block_sum(double (&x)[N], double (*red)[kGuardThreads])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < N; ++q) red[q][t] = x[q];
    __syncthreads();
    for (int h = kGuardThreads / 2; h > 0; h >>= 1) {
        if (t < h)
#pragma unroll
            for (int q = 0; q < N; ++q) red[q][t] = red[q][t] + red[q][t + h];
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < N; ++q) x[q] = red[q][0];
    __syncthreads();
}

__global__ void k_guard_clear(unsigned *counters)
{
    if (threadIdx.x < 3) counters[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(kGuardThreads) void k_peak_guard(
    const PairDesc *__restrict__ desc, int dtype, int C, int W, int E, int rm, int exclude,
    float tau, const float *__restrict__ gamma, const int *__restrict__ near,
    int2 *__restrict__ loc, int2 *__restrict__ window, float *__restrict__ peak,
    float *__restrict__ snr, unsigned *__restrict__ counters)
{
    __shared__ int cand[kGuardMax];
    __shared__ int ncand;
    __shared__ double red[5][kGuardThreads];
    const int pair = blockIdx.x;
    // the peak kernel counted the cells within tau of the peak
    if (near[pair] <= 1) return;
    const int t = threadIdx.x;
    const int P = W - C + 1, cells = P * P;
    const float *g = gamma + (size_t)pair * cells;
    const float pk = peak[pair];
    if (pk != pk) return;  // no peak (all NaN): nothing to guard
    if (t == 0) ncand = 0;
    __syncthreads();
    const float floor_ = pk - tau;
    for (int k = t; k < cells; k += kGuardThreads) {
        const float x = g[k];
        if (x >= floor_) {
            const int at = atomicAdd(&ncand, 1);
            if (at < kGuardMax) cand[at] = k;
        }
    }
    __syncthreads();
    const int n = ncand;
    if (n <= 1) return;
    if (n > kGuardMax) {
        if (t == 0) atomicAdd(&counters[2], 1u);
        return;  // keep the FP8 peak
    }
    if (t == 0) {
        // row-major order, so ties keep the first maximum
        for (int a = 1; a < n; ++a) {
            const int key = cand[a];
            int b = a - 1;
            while (b >= 0 && cand[b] > key) {
                cand[b + 1] = cand[b];
                --b;
            }
            cand[b + 1] = key;
        }
        atomicAdd(&counters[0], 1u);
    }
    __syncthreads();
    const PairDesc d = desc[pair];
    const double nn = (double)C * C;
    double best = -INFINITY;
    int bidx = cand[0];
    for (int c = 0; c < n; ++c) {
        const int k = cand[c];
        const int u = k / P, v = k - u * P;
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // r, r^2, t, t^2, r t
        for (int e = t; e < C * C; e += kGuardThreads) {
            const int i = e / C, j = e - i * C;
            const double r = detect64(dtype, d.ref, (size_t)(d.ref_line + i) * d.ref_stride +
                                                        d.ref_sample + j);
            const double x = detect64(dtype, d.tgt, (size_t)(d.tgt_line + u + i) * d.tgt_stride +
                                                        d.tgt_sample + v + j);
            acc[0] += r;
            acc[1] += r * r;
            acc[2] += x;
            acc[3] += x * x;
            acc[4] += r * x;
        }
        block_sum(acc, red);
        const double vr = acc[1] - acc[0] * acc[0] / nn, vt = acc[3] - acc[2] * acc[2] / nn;
        const double gc = (vr > 0.0 && vt > 0.0) ? (acc[4] - acc[0] * acc[2] / nn) / sqrt(vr * vt)
                                                 : -INFINITY;
        if (gc > best) {
            best = gc;
            bidx = k;
        }
    }
    const int2 old = loc[pair];
    const int row = bidx / P, col = bidx - row * P;
    if (row == old.x && col == old.y) return;
    // the new peak: nudged window, the surface's value, SNR around it
    double acc[2] = {0.0, 0.0};
    for (int k = t; k < cells; k += kGuardThreads) {
        const float x = g[k];
        const int u = k / P, v = k - u * P;
        if (x != x || (abs(u - row) <= exclude && abs(v - col) <= exclude)) continue;
        acc[0] += (double)x * (double)x;
        acc[1] += 1.0;
    }
    block_sum(acc, red);
    if (t == 0) {
        int left = col - rm;
        if (left < 0) left = 0;
        if (left + E > W) left = W - E;
        int top = row - rm;
        if (top < 0) top = 0;
        if (top + E > W) top = W - E;
        const float npk = g[bidx];
        const double p2 = (double)npk * (double)npk;
        loc[pair] = make_int2(row, col);
        window[pair] = make_int2(top, left);
        peak[pair] = npk;
        snr[pair] = (acc[1] == 0.0 || acc[0] <= 0.0) ? INFINITY : (float)(p2 / (acc[0] / acc[1]));
        atomicAdd(&counters[1], 1u);
    }
}

// ---------------------------------------------------------------------------
// band-limited upsampling y = U x U^T of complex N x N tiles on the matrix
// cores (FP32 MFMA 16x16x4).  U is real, M x N with M = rf N, stored padded to
// Mp x Np (multiples of 16, zero-filled).  One workgroup per pair; each wave
// owns 16-row strips of the output: T = U[strip, :] x (from the raster, both
// planes) into LDS, then y[strip, :] = T U^T, detected on the way out.
//
// MFMA 16x16x4 f32 operand layout (lane l): A[m = l % 16][k = l / 16],
// B[k = l / 16][n = l % 16], D[m = 4 (l / 16) + i][n = l % 16], i = 0..3.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int kUpWaves = 4;

template <int DTYPE>
__global__ __launch_bounds__(kUpWaves * kWave) void k_upsample(
    const PairDesc *__restrict__ desc, int role, const int2 *__restrict__ window, int N, int Np,
    int M, int Mp, const float *__restrict__ U, float *__restrict__ amp)
{
    extern __shared__ float tstrip[];  // [wave][plane][16][Np]
    const int pair = blockIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const PairDesc d = desc[pair];
    const void *base = role == 0 ? d.ref : d.tgt;
    const size_t stride = role == 0 ? d.ref_stride : d.tgt_stride;
    int line = role == 0 ? d.ref_line : d.tgt_line;
    int sample = role == 0 ? d.ref_sample : d.tgt_sample;
    if (role == 1) {
        const int2 tl = window[pair];
        line += tl.x;
        sample += tl.y;
    }
    float *Tr = tstrip + (size_t)wave * 2 * 16 * Np;
    float *Ti = Tr + 16 * Np;
    float *out = amp + (size_t)pair * M * M;

    for (int strip = wave; strip < Mp / 16; strip += kUpWaves) {
        const int p0 = 16 * strip;
        // stage 1: T (16 x Np) = U[p0 : p0 + 16, :] x
        for (int nt = 0; nt < Np / 16; ++nt) {
            floatx4 accR = {0.0f, 0.0f, 0.0f, 0.0f}, accI = {0.0f, 0.0f, 0.0f, 0.0f};
            const int n = nt * 16 + lm;
            for (int k0 = 0; k0 < Np; k0 += 4) {
                const int k = k0 + lk;
                const float a = U[(size_t)(p0 + lm) * Np + k];
                float br = 0.0f, bi = 0.0f;
                if (k < N && n < N) {
                    const size_t off = (size_t)(line + k) * stride + sample + n;
                    if (DTYPE == 0) {
                        const float2 z = static_cast<const float2 *>(base)[off];
                        br = z.x;
                        bi = z.y;
                    } else {
                        br = static_cast<const float *>(base)[off];
                    }
                }
                accR = __builtin_amdgcn_mfma_f32_16x16x4f32(a, br, accR, 0, 0, 0);
                if (DTYPE == 0) accI = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bi, accI, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                Tr[(4 * lk + i) * Np + n] = accR[i];
                Ti[(4 * lk + i) * Np + n] = (DTYPE == 0) ? accI[i] : 0.0f;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // stage 2: y[p0 : p0 + 16, :] = T U^T
        for (int qt = 0; qt < Mp / 16; ++qt) {
            floatx4 accR = {0.0f, 0.0f, 0.0f, 0.0f}, accI = {0.0f, 0.0f, 0.0f, 0.0f};
            const int q = qt * 16 + lm;
            for (int k0 = 0; k0 < Np; k0 += 4) {
                const int k = k0 + lk;
                const float b = U[(size_t)q * Np + k];
                accR = __builtin_amdgcn_mfma_f32_16x16x4f32(Tr[lm * Np + k], b, accR, 0, 0, 0);
                accI = __builtin_amdgcn_mfma_f32_16x16x4f32(Ti[lm * Np + k], b, accI, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int p = p0 + 4 * lk + i;
                if (p < M && q < M) {
                    const double re = accR[i], im = accI[i];
                    out[(size_t)p * M + q] = (float)sqrt(re * re + im * im);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// x2 upsampling with the tile resident in LDS (N <= 96).  With rf = 2 the
// refined rows 2a are the coarse rows and rows 2a + 1 are a half-sample shift,
// a circulant H[a][n] = g((a - n) mod N), so
//   y[2a][2b] = x,  y[2a+1][2b] = (H x),  y[2a+s][2b+1] = (T_s H^T),
// with T_0 = x, T_1 = H x: three N^3 products instead of rf^3 N^3 / ... of
// the dense form.  8 waves share the output tiles.  The circulant is read
// from g laid out periodically in LDS (H[r][c] = gx[r - c + Np]: fixed offsets
// in the MFMA loops), the tile is gathered with kUp2Batch loads in flight
// per thread, and all four phases of an output 2 x 2 block are written
// together at the end (paired stores).
// ---------------------------------------------------------------------------
constexpr int kUp2Waves = 8;
constexpr int kUp2MaxN = 96;
constexpr int kUp2Batch = 8;

// |z| in FP32 (v_sqrt_f32, within 1 ulp): the refined stage is compared with
// the FP64 oracle by tolerance
__device__ inline float amplitude(float re, float im)
{
    return __builtin_amdgcn_sqrtf(re * re + im * im);
}

__device__ inline uint16_t quantize(int qmode, float x)
{
    if (qmode == kUpFP16) return __builtin_bit_cast(uint16_t, (_Float16)x);
    const uint32_t b = __float_as_uint(x);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

__device__ inline float dequantize(int qmode, uint16_t h)
{
    if (qmode == kUpFP16) return (float)__builtin_bit_cast(_Float16, h);
    return __uint_as_float((uint32_t)h << 16);
}

// power of two bringing max-abs into [0.5, 1) for FP16 and into [128, 256)
// for FP8 (e4m3 tops out at 448); 1 for BF16, whose range is FP32's
__device__ inline double quant_scale(int qmode, float maxabs)
{
    if ((qmode != kUpFP16 && qmode != kUpFP8) || !(maxabs > 0.0f)) return 1.0;
    int e;
    frexp((double)maxabs, &e);
    return ldexp(1.0, qmode == kUpFP8 ? 8 - e : -e);
}


// sum (FP64) and max over the workgroup; every thread gets the results
__device__ inline void block_sum_max(double *red, float *redm, double &sum, float &mx)
{
    for (int d = kWave / 2; d > 0; d >>= 1) {
        sum += __shfl_xor(sum, d, kWave);
        mx = fmaxf(mx, __shfl_xor(mx, d, kWave));
    }
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[wave] = sum;
        redm[wave] = mx;
    }
    __syncthreads();
    sum = 0.0;
    mx = 0.0f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        sum += red[w];
        mx = fmaxf(mx, redm[w]);
    }
}

template <int DTYPE, int QMODE>
__global__ __launch_bounds__(kUp2Waves * kWave) void k_upsample2(
    const PairDesc *__restrict__ desc, int role, const int2 *__restrict__ window, int N, int Np,
    const float *__restrict__ g, float *__restrict__ amp, uint16_t *__restrict__ qout,
    uint16_t *__restrict__ qlo, int cs, MixedStats *__restrict__ stats)
{
    extern __shared__ float lds[];
    __shared__ double red[kUp2Waves], red2[kUp2Waves];
    __shared__ float redm[kUp2Waves];
    const int ls = Np + 4;
    float *xr = lds, *xi = xr + Np * ls, *tr = xi + Np * ls, *ti = tr + Np * ls;
    float *gx = ti + Np * ls;   // gx[j] = g((j - Np) mod N), j < 2 Np
    const int pair = blockIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const PairDesc d = desc[pair];
    const void *base = role == 0 ? d.ref : d.tgt;
    const size_t stride = role == 0 ? d.ref_stride : d.tgt_stride;
    int line = role == 0 ? d.ref_line : d.tgt_line;
    int sample = role == 0 ? d.ref_sample : d.tgt_sample;
    if (role == 1) {
        const int2 tl = window[pair];
        line += tl.x;
        sample += tl.y;
    }
    double asum = 0.0;
    float amax = 0.0f;
    const int total = Np * Np;
    for (int k0 = threadIdx.x; k0 < total; k0 += kUp2Batch * blockDim.x) {
        float re[kUp2Batch], im[kUp2Batch];
#pragma unroll
        for (int j = 0; j < kUp2Batch; ++j) {
            const int k = k0 + j * blockDim.x;
            const int r = k / Np, c = k - r * Np;
            re[j] = 0.0f;
            im[j] = 0.0f;
            if (k < total && r < N && c < N) {
                const size_t off = (size_t)(line + r) * stride + sample + c;
                if (DTYPE == 0) {
                    const float2 z = static_cast<const float2 *>(base)[off];
                    re[j] = z.x;
                    im[j] = z.y;
                } else {
                    re[j] = static_cast<const float *>(base)[off];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kUp2Batch; ++j) {
            const int k = k0 + j * blockDim.x;
            if (k < total) {
                const int r = k / Np, c = k - r * Np;
                xr[r * ls + c] = re[j];
                xi[r * ls + c] = im[j];
                if (QMODE != kUpFP32) {
                    const float a = amplitude(re[j], im[j]);
                    asum += a;
                    amax = fmaxf(amax, a);
                }
            }
        }
    }
    for (int k = threadIdx.x; k < 2 * Np; k += blockDim.x) {
        const int t = (k - Np) % N;
        gx[k] = g[t < 0 ? t + N : t];
    }
    const int M = 2 * N;
    float qc = 0.0f, qs = 1.0f;
    if (QMODE != kUpFP32) {
        // the quantization constants come from the coarse tile: its mean
        // amplitude is removed, and twice its peak bounds the upsampled values
        block_sum_max(red, redm, asum, amax);
        qc = (float)(asum / ((double)N * N));
        qs = (float)quant_scale(QMODE, 2.0f * amax);
        // zero padding of the chip rows
        if (role == 0)
            for (int k = threadIdx.x; k < M * (cs - M); k += blockDim.x) {
                const int p = k / (cs - M), c = k - p * (cs - M);
                const int col = c < 16 ? c : M + c;
                qout[((size_t)pair * M + p) * cs + col] = 0;
                if (qlo) qlo[((size_t)pair * M + p) * cs + col] = 0;
            }
    }
    __syncthreads();

    const int tiles = Np / 16;
    // stage 1: T_1 = H x; a wave takes two output tiles of one tile row at a
    // time, four independent accumulation chains
    for (int pr = wave; pr < tiles * ((tiles + 1) / 2); pr += kUp2Waves) {
        const int mt = pr / ((tiles + 1) / 2);
        const int nt0 = 2 * (pr - mt * ((tiles + 1) / 2));
        const bool two = nt0 + 1 < tiles;
        // H[m][k] for rows m >= N or columns k >= N multiplies zeros or
        // lands in outputs that are not written
        const float *hrow = gx + (16 * mt + lm - lk + Np);
        const float *xrow = xr + lk * ls + 16 * nt0 + lm;
        const float *irow = xi + lk * ls + 16 * nt0 + lm;
        floatx4 aR0 = {0.0f, 0.0f, 0.0f, 0.0f}, aI0 = aR0, aR1 = aR0, aI1 = aR0;
        for (int k16 = 0; k16 < Np; k16 += 16)
#pragma unroll
            for (int k0 = k16; k0 < k16 + 16; k0 += 4) {
                const float a = hrow[-k0];
                aR0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xrow[k0 * ls], aR0, 0, 0, 0);
                if (DTYPE == 0)
                    aI0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, irow[k0 * ls], aI0, 0, 0, 0);
                if (two) {
                    aR1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xrow[k0 * ls + 16], aR1, 0, 0, 0);
                    if (DTYPE == 0)
                        aI1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, irow[k0 * ls + 16], aI1, 0,
                                                                   0, 0);
                }
            }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = (16 * mt + 4 * lk + i) * ls + 16 * nt0 + lm;
            tr[row] = aR0[i];
            ti[row] = (DTYPE == 0) ? aI0[i] : 0.0f;
            if (two) {
                tr[row + 16] = aR1[i];
                ti[row + 16] = (DTYPE == 0) ? aI1[i] : 0.0f;
            }
        }
    }
    __syncthreads();

    float *out = amp ? amp + (size_t)pair * M * M : nullptr;
    double s1 = 0.0, s2 = 0.0;
    // y[p][q], y[p][q + 1] (q even)
    auto emit2 = [&](int p, int q, float v0, float v1) {
        if (QMODE == kUpFP32 || role == 1)
            *reinterpret_cast<float2 *>(out + (size_t)p * M + q) = make_float2(v0, v1);
        if (QMODE != kUpFP32) {
            const float x0 = (v0 - qc) * qs, x1 = (v1 - qc) * qs;
            const uint16_t h0 = quantize(QMODE, x0), h1 = quantize(QMODE, x1);
            // split form: the rounding remainders as a second operand
            uint16_t l0 = 0, l1 = 0;
            if (qlo) {
                l0 = quantize(QMODE, x0 - dequantize(QMODE, h0));
                l1 = quantize(QMODE, x1 - dequantize(QMODE, h1));
            }
            const size_t at = role == 0 ? ((size_t)pair * M + p) * cs + 16 + q
                                        : ((size_t)pair * M + p) * M + q;
            *reinterpret_cast<uint32_t *>(qout + at) = (uint32_t)h0 | ((uint32_t)h1 << 16);
            if (qlo) *reinterpret_cast<uint32_t *>(qlo + at) = (uint32_t)l0 | ((uint32_t)l1 << 16);
            if (role == 0) {
                const double b0 = ((double)dequantize(QMODE, h0) +
                                   (qlo ? (double)dequantize(QMODE, l0) : 0.0)) / qs;
                const double b1 = ((double)dequantize(QMODE, h1) +
                                   (qlo ? (double)dequantize(QMODE, l1) : 0.0)) / qs;
                s1 += b0 + b1;
                s2 += b0 * b0 + b1 * b1;
            }
        }
    };
    // stage 2: odd columns, y[2a + s][2b + 1] = (T_s H^T)[a][b]; a wave takes
    // the same output tile of both phases s at a time (four chains, shared B),
    // and writes the 2 x 2 output blocks of its tile
    for (int tile = wave; tile < tiles * tiles; tile += kUp2Waves) {
        const int mt = tile / tiles, nt = tile - mt * tiles;
        const int n = 16 * nt + lm;
        const float *hrow = gx + (n - lk + Np);
        const int arow = (16 * mt + lm) * ls + lk;
        floatx4 xR = {0.0f, 0.0f, 0.0f, 0.0f}, xI = xR, tR = xR, tI = xR;
        for (int k16 = 0; k16 < Np; k16 += 16)
#pragma unroll
            for (int k0 = k16; k0 < k16 + 16; k0 += 4) {
                const float b = hrow[-k0];
                const int ai = arow + k0;
                xR = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[ai], b, xR, 0, 0, 0);
                xI = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[ai], b, xI, 0, 0, 0);
                tR = __builtin_amdgcn_mfma_f32_16x16x4f32(tr[ai], b, tR, 0, 0, 0);
                tI = __builtin_amdgcn_mfma_f32_16x16x4f32(ti[ai], b, tI, 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int a = 16 * mt + 4 * lk + i;
            if (a < N && n < N) {
                const int at = a * ls + n;
                emit2(2 * a, 2 * n, amplitude(xr[at], xi[at]), amplitude(xR[i], xI[i]));
                emit2(2 * a + 1, 2 * n, amplitude(tr[at], ti[at]), amplitude(tR[i], tI[i]));
            }
        }
    }
    if (QMODE != kUpFP32) {
        if (role == 0) {
            float unused = 0.0f;
            block_sum_max(red, redm, s1, unused);
            block_sum_max(red2, redm, s2, unused);
            if (threadIdx.x == 0) {
                stats[pair].sumr = s1;
                stats[pair].sumr2 = s2;
                stats[pair].scale_r = qs;
            }
        } else if (threadIdx.x == 0) {
            stats[pair].scale_t = qs;
            stats[pair].mean_t = qc;
        }
    }
}

// ---------------------------------------------------------------------------
// x2 upsampling for the BF16/FP16/FP8 forms: the same three circulant
// products as k_upsample2, on v_mfma_f32_16x16x32_f16 with FP32 accumulation.
// The tile is read once into registers (its peak and mean amplitude taken
// there) and scaled by a power of two into FP16 range; LDS holds it twice,
// row- and column-major, because the first product contracts over its rows
// and the second over its columns (complex tiles: four planes, 78 KiB for
// N = 96, so two workgroups share a CU; real tiles: two planes, up to
// N = 160).  Only the tiles that hold samples are multiplied.  The circulant comes as
// precomputed FP16 fragments (hfrag, one 16 x 32 block per 64 lanes x 8
// values: H[16 mt + l % 16][32 kc + 8 (l / 16) + e]), read from global
// memory (L2-resident, shared by every pair), which serve as the A operand
// of T = H x and the B operand of Y_s = T_s H^T.  T = H x is kept in FP16 in
// the space of the column-major copy.  Rounding x and T to FP16 (11 bits)
// stays well below the BF16/FP16 rounding of the operands written for the
// contraction.
// ---------------------------------------------------------------------------
constexpr int kUp2hMaxN = 96;       // complex tiles
constexpr int kUp2hMaxNReal = 160;  // real tiles: two planes instead of four

// largest tile edge (padded to 32) k_upsample2h takes for the raster type
__host__ __device__ constexpr int up2h_max_n(int dtype) { return dtype == 1 ? kUp2hMaxNReal : kUp2hMaxN; }

typedef _Float16 halfx8_t __attribute__((ext_vector_type(8)));
typedef _Float16 halfx2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float floatx2_t __attribute__((ext_vector_type(2)));

__device__ inline floatx4 mfma_f16(uint4 a, uint4 b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(halfx8_t, a),
                                                  __builtin_bit_cast(halfx8_t, b), c, 0, 0, 0);
}

// LDS of k_upsample2h (bytes): four FP16 planes [Np][Np + 8], two for real tiles
__host__ __device__ inline size_t up2h_plane_bytes(int Np) { return (size_t)Np * (Np + 8) * 2; }
__host__ __device__ inline size_t up2h_lds_bytes(int Np, int dtype)
{
    return (dtype == 1 ? 2 : 4) * up2h_plane_bytes(Np);
}

// MAXN: the largest padded tile edge of the instantiation; the register arrays
// are sized for it, so a smaller MAXN (64, for 64-sample chips) leaves
// room for more workgroups per CU
template <int DTYPE, int QMODE, int MAXN>
__global__ __launch_bounds__(kUp2Waves * kWave) void k_upsample2h(
    const PairDesc *__restrict__ desc, int role, const int2 *__restrict__ window, int N, int Np,
    const uint4 *__restrict__ hfrag, float *__restrict__ amp, uint16_t *__restrict__ qout, int cs,
    MixedStats *__restrict__ stats)
{
    // real tiles (DTYPE 1) have no imaginary planes, and may be larger
    constexpr bool kReal = DTYPE == 1;
    constexpr int kMaxN = MAXN;
    static_assert(MAXN % 32 == 0 && MAXN <= up2h_max_n(DTYPE), "tile edge of the instantiation");
    extern __shared__ uint4 lds4[];
    __shared__ double red[kUp2Waves], red2[kUp2Waves];
    __shared__ float redm[kUp2Waves];
    unsigned char *lb = reinterpret_cast<unsigned char *>(lds4);
    const int Ls = Np + 8;
    const size_t pb = up2h_plane_bytes(Np);
    _Float16 *xr = reinterpret_cast<_Float16 *>(lb);                         // x, row-major
    _Float16 *cr = reinterpret_cast<_Float16 *>(lb + pb);                    // x column-major, then T row-major
    _Float16 *xi = kReal ? nullptr : reinterpret_cast<_Float16 *>(lb + 2 * pb);
    _Float16 *ci = kReal ? nullptr : reinterpret_cast<_Float16 *>(lb + 3 * pb);
    const int MT = Np / 16, KC = Np / 32;
    const int pair = blockIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const PairDesc d = desc[pair];
    const void *base = role == 0 ? d.ref : d.tgt;
    const size_t stride = role == 0 ? d.ref_stride : d.tgt_stride;
    int line = role == 0 ? d.ref_line : d.tgt_line;
    int sample = role == 0 ? d.ref_sample : d.tgt_sample;
    if (role == 1) {
        const int2 tl = window[pair];
        line += tl.x;
        sample += tl.y;
    }
    auto load = [&](int r, int c, float &re, float &im) {
        const size_t off = (size_t)(line + r) * stride + sample + c;
        if (DTYPE == 0) {
            const float2 z = static_cast<const float2 *>(base)[off];
            re = z.x;
            im = z.y;
        } else {
            re = static_cast<const float *>(base)[off];
            im = 0.0f;
        }
    };

    // the tile in 2 x 2 blocks, read once into registers: a wave takes 8 x 8
    // blocks at a time (block row lane / 8, block column lane % 8), so each
    // row segment a wave reads is 128 contiguous bytes and the row- and
    // column-major LDS writes below hit 64 different banks
    constexpr int kMaxJ = (kMaxN / 16) * (kMaxN / 16) / kUp2Waves + 1;
    const int TB = Np / 16, ntb = TB * TB;
    const int rsub = 2 * (lane >> 3), csub = 2 * (lane & 7);
    float bre[kMaxJ][4], bim[kMaxJ][4];
#pragma unroll
    for (int j = 0; j < kMaxJ; ++j) {
        const int tb = wave + kUp2Waves * j;
        const int r = 16 * (tb / TB) + rsub, c = 16 * (tb % TB) + csub;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int rr = r + (e >> 1), cc = c + (e & 1);
            bre[j][e] = bim[j][e] = 0.0f;
            if (tb < ntb && rr < N && cc < N) load(rr, cc, bre[j][e], bim[j][e]);
        }
    }
    // the sum and peak of the coarse amplitudes (padding adds zeros)
    double asum = 0.0;
    float amax = 0.0f;
#pragma unroll
    for (int j = 0; j < kMaxJ; ++j) {
        float part = 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float a = amplitude(bre[j][e], bim[j][e]);
            part += a;
            amax = fmaxf(amax, a);
        }
        asum += (double)part;
    }
    block_sum_max(red, redm, asum, amax);
    const int M = 2 * N;
    const float qc = (float)(asum / ((double)N * N));
    const float qs = (float)quant_scale(QMODE, 2.0f * amax);
    // FP16 range: |re|, |im| <= amax, scaled below 2^13 (T = H x stays below 2^15)
    int ex = 0;
    if (amax > 0.0f) frexpf(amax, &ex);
    const float sx = amax > 0.0f ? ldexpf(1.0f, 13 - ex) : 1.0f, inv_sx = 1.0f / sx;
    if (role == 0) {
        // zero padding of the chip rows: columns [0, 16) and [M + 16, cs), in
        // 32-bit words, 32 threads per row
        const int pw = (cs - M) / 2, sub = threadIdx.x & 31;
        if (sub < pw)
            for (int p = threadIdx.x >> 5; p < M; p += (int)blockDim.x >> 5) {
                const int col = sub < 8 ? 2 * sub : M + 2 * sub;
                *reinterpret_cast<uint32_t *>(qout + ((size_t)pair * M + p) * cs + col) = 0u;
            }
    }
    // the scaled FP16 tile, zero-padded to Np x Np, row- and column-major
    // (pairs of values as 32-bit words)
    auto pack = [](float lo, float hi) {
        const halfx2_t h = __builtin_convertvector((floatx2_t){lo, hi}, halfx2_t);
        return __builtin_bit_cast(uint32_t, h);
    };
#pragma unroll
    for (int j = 0; j < kMaxJ; ++j) {
        const int tb = wave + kUp2Waves * j;
        if (tb >= ntb) continue;
        const int r = 16 * (tb / TB) + rsub, c = 16 * (tb % TB) + csub;
        const float *v = bre[j], *w = bim[j];
        *reinterpret_cast<uint32_t *>(xr + r * Ls + c) = pack(v[0] * sx, v[1] * sx);
        *reinterpret_cast<uint32_t *>(xr + (r + 1) * Ls + c) = pack(v[2] * sx, v[3] * sx);
        *reinterpret_cast<uint32_t *>(cr + c * Ls + r) = pack(v[0] * sx, v[2] * sx);
        *reinterpret_cast<uint32_t *>(cr + (c + 1) * Ls + r) = pack(v[1] * sx, v[3] * sx);
        if constexpr (!kReal) {
            *reinterpret_cast<uint32_t *>(xi + r * Ls + c) = pack(w[0] * sx, w[1] * sx);
            *reinterpret_cast<uint32_t *>(xi + (r + 1) * Ls + c) = pack(w[2] * sx, w[3] * sx);
            *reinterpret_cast<uint32_t *>(ci + c * Ls + r) = pack(w[0] * sx, w[2] * sx);
            *reinterpret_cast<uint32_t *>(ci + (c + 1) * Ls + r) = pack(w[1] * sx, w[3] * sx);
        }
    }
    __syncthreads();

    // stage 1: T = H x, tiles (mt, ct); a wave keeps its tiles until every
    // wave is done with the column-major copy.  Only the NT x NT tiles that
    // hold samples are computed: the rest of that copy (x[n][a] with n >= N)
    // is zero, which is what T holds there and what the products below read
    constexpr int kMaxT = (kMaxN / 16) * (kMaxN / 16) / kUp2Waves + 1;
    floatx4 tR[kMaxT], tI[kMaxT];
    const int NT = (N + 15) / 16, ntile = NT * NT;
#pragma unroll
    for (int q = 0; q < kMaxT; ++q) {
        tR[q] = tI[q] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        const int tile = wave + kUp2Waves * q;
        if (tile >= ntile) continue;
        const int mt = tile / NT, ct = tile - mt * NT;
        const _Float16 *br = cr + (16 * ct + lm) * Ls + 8 * lk;
        const _Float16 *bi = kReal ? br : ci + (16 * ct + lm) * Ls + 8 * lk;
        uint4 fa[kMaxN / 32];
#pragma unroll
        for (int kc = 0; kc < kMaxN / 32; ++kc)
            if (kc < KC) fa[kc] = hfrag[(mt * KC + kc) * kWave + lane];
#pragma unroll
        for (int kc = 0; kc < kMaxN / 32; ++kc) {
            if (kc >= KC) break;
            const uint4 a = fa[kc];
            tR[q] = mfma_f16(a, *reinterpret_cast<const uint4 *>(br + 32 * kc), tR[q]);
            if constexpr (!kReal)
                tI[q] = mfma_f16(a, *reinterpret_cast<const uint4 *>(bi + 32 * kc), tI[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxT; ++q) {
        const int tile = wave + kUp2Waves * q;
        if (tile >= ntile) continue;
        const int mt = tile / NT, ct = tile - mt * NT;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int at = (16 * mt + 4 * lk + i) * Ls + 16 * ct + lm;
            cr[at] = (_Float16)tR[q][i];
            if constexpr (!kReal) ci[at] = (_Float16)tI[q][i];
        }
    }
    __syncthreads();

    float *out = (role == 1 && amp) ? amp + (size_t)pair * M * M : nullptr;
    uint16_t *qbase = qout + (size_t)pair * M * (role == 0 ? cs : M) + (role == 0 ? 16 : 0);
    const int qstride = role == 0 ? cs : M;
    // (a / sx - qc) qs == fma(a, qs / sx, -qc qs): sx and qs are powers of two
    const float ka = inv_sx * qs, kb = -qc * qs;
    // chip sums in the operand's units (exact: qs is a power of two), scaled back at the end
    double s1 = 0.0, s2 = 0.0;
    // y[p][q], y[p][q + 1] (q even) from amplitudes in the tile's FP16 scale
    auto emit2 = [&](int p, int q, float a0, float a1) {
        if (out)
            *reinterpret_cast<float2 *>(out + (p * M + q)) = make_float2(a0 * inv_sx, a1 * inv_sx);
        const floatx2_t x = {fmaf(a0, ka, kb), fmaf(a1, ka, kb)};
        uint32_t h;
        if constexpr (QMODE == kUpFP16)
            h = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, halfx2_t));
        else
            h = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2_t));
        *reinterpret_cast<uint32_t *>(qbase + (p * qstride + q)) = h;
        if (role == 0) {
            const double b0 = (double)dequantize(QMODE, (uint16_t)h);
            const double b1 = (double)dequantize(QMODE, (uint16_t)(h >> 16));
            s1 += b0 + b1;
            s2 += b0 * b0 + b1 * b1;
        }
    };
    // stage 2: odd columns, Y_s = T_s H^T for T_0 = x and T_1 = T, then the
    // four phases of each output 2 x 2 block; each tile's circulant fragments
    // are loaded during the tile before it
    uint4 fnext[kMaxN / 32];
    auto load_frag = [&](int tile) {
        const int bt = tile % NT;
#pragma unroll
        for (int kc = 0; kc < kMaxN / 32; ++kc)
            if (kc < KC) fnext[kc] = hfrag[(bt * KC + kc) * kWave + lane];
    };
    if (wave < ntile) load_frag(wave);
    for (int tile = wave; tile < ntile; tile += kUp2Waves) {
        const int at = tile / NT, bt = tile - at * NT;
        const int arow = (16 * at + lm) * Ls + 8 * lk;
        floatx4 xR = {0.0f, 0.0f, 0.0f, 0.0f}, xI = xR, yR = xR, yI = xR;
        uint4 fb[kMaxN / 32];
#pragma unroll
        for (int kc = 0; kc < kMaxN / 32; ++kc) fb[kc] = fnext[kc];
        if (tile + kUp2Waves < ntile) load_frag(tile + kUp2Waves);
#pragma unroll
        for (int kc = 0; kc < kMaxN / 32; ++kc) {
            if (kc >= KC) break;
            const uint4 b = fb[kc];
            const int o = arow + 32 * kc;
            xR = mfma_f16(*reinterpret_cast<const uint4 *>(xr + o), b, xR);
            yR = mfma_f16(*reinterpret_cast<const uint4 *>(cr + o), b, yR);
            if constexpr (!kReal) {
                xI = mfma_f16(*reinterpret_cast<const uint4 *>(xi + o), b, xI);
                yI = mfma_f16(*reinterpret_cast<const uint4 *>(ci + o), b, yI);
            }
        }
        const int bcol = 16 * bt + lm;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int a = 16 * at + 4 * lk + i;
            if (a < N && bcol < N) {
                const int t = a * Ls + bcol;
                const float xit = kReal ? 0.0f : (float)xi[t], cit = kReal ? 0.0f : (float)ci[t];
                emit2(2 * a, 2 * bcol, amplitude((float)xr[t], xit), amplitude(xR[i], xI[i]));
                emit2(2 * a + 1, 2 * bcol, amplitude((float)cr[t], cit), amplitude(yR[i], yI[i]));
            }
        }
    }
    if (role == 0) {
        float unused = 0.0f;
        block_sum_max(red, redm, s1, unused);
        block_sum_max(red2, redm, s2, unused);
        if (threadIdx.x == 0) {
            stats[pair].sumr = s1 / qs;
            stats[pair].sumr2 = s2 / ((double)qs * qs);
            stats[pair].scale_r = qs;
        }
    } else if (threadIdx.x == 0) {
        stats[pair].scale_t = qs;
        stats[pair].mean_t = qc;
    }
}

// ---------------------------------------------------------------------------
// x2 upsampling of larger tiles (96 < N <= kUpBandMaxN) in two passes over
// bands of 16, on FP32 matrix cores, with the circulant H[a][n] = g((a - n)
// mod N) read from a periodic copy of g in LDS (gx[Np + a - n]):
//   pass 1 (column bands): T = H x for 16 columns, to global memory, and the
//          band's amplitude sum and peak;
//   pass 2 (row bands): Y_s = T_s H^T for 16 rows of x and of T, and the four
//          phases of output rows 2a, 2a + 1, written whole -- FP32 amplitudes
//          and, in the quantized modes, the contraction's operand with the
//          constants of all column bands (summed in band order) and the
//          band's share of the chip statistics;
//   k_up2_finish sums the chip statistics of the row bands in order.
// Small LDS footprints (two 16-wide strips) keep many workgroups per CU.
// ---------------------------------------------------------------------------
constexpr int kUpBandThreads = 256;
constexpr int kUpBandMaxN = 512;

template <int DTYPE>
__device__ inline void load_sample(const void *base, size_t off, float &re, float &im)
{
    if (DTYPE == 0) {
        const float2 z = static_cast<const float2 *>(base)[off];
        re = z.x;
        im = z.y;
    } else {
        re = static_cast<const float *>(base)[off];
        im = 0.0f;
    }
}

__device__ inline void tile_origin(const PairDesc &d, int role, const int2 *window, int pair,
                                   const void *&base, size_t &stride, int &line, int &sample)
{
    base = role == 0 ? d.ref : d.tgt;
    stride = role == 0 ? d.ref_stride : d.tgt_stride;
    line = role == 0 ? d.ref_line : d.tgt_line;
    sample = role == 0 ? d.ref_sample : d.tgt_sample;
    if (role == 1) {
        const int2 tl = window[pair];
        line += tl.x;
        sample += tl.y;
    }
}

template <int DTYPE>
__global__ __launch_bounds__(kUpBandThreads) void k_up2_cols(
    const PairDesc *__restrict__ desc, int role, const int2 *__restrict__ window, int N, int Np,
    const float *__restrict__ g, float2 *__restrict__ T, double2 *__restrict__ band)
{
    extern __shared__ float ub[];
    float *xr = ub, *xi = xr + Np * 16, *gx = xi + Np * 16;   // the band [Np][16], g
    __shared__ double red[kUpBandThreads / kWave];
    __shared__ float redm[kUpBandThreads / kWave];
    const int pair = blockIdx.x, cb = blockIdx.y, nb = gridDim.y, c0 = 16 * cb;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const PairDesc d = desc[pair];
    const void *base;
    size_t stride;
    int line, sample;
    tile_origin(d, role, window, pair, base, stride, line, sample);
    double asum = 0.0;
    float amax = 0.0f;
    const int total = Np * 16;
    for (int k0 = threadIdx.x; k0 < total; k0 += kUp2Batch * blockDim.x) {
        float re[kUp2Batch], im[kUp2Batch];
#pragma unroll
        for (int j = 0; j < kUp2Batch; ++j) {
            const int k = k0 + j * blockDim.x;
            const int r = k >> 4, c = c0 + (k & 15);
            re[j] = im[j] = 0.0f;
            if (k < total && r < N && c < N)
                load_sample<DTYPE>(base, (size_t)(line + r) * stride + sample + c, re[j], im[j]);
        }
#pragma unroll
        for (int j = 0; j < kUp2Batch; ++j) {
            const int k = k0 + j * blockDim.x;
            if (k < total) {
                xr[k] = re[j];
                xi[k] = im[j];
                const float a = amplitude(re[j], im[j]);   // zero outside the tile
                asum += a;
                amax = fmaxf(amax, a);
            }
        }
    }
    for (int k = threadIdx.x; k < 2 * Np; k += blockDim.x) {
        const int t = (k - Np) % N;
        gx[k] = g[t < 0 ? t + N : t];
    }
    block_sum_max(red, redm, asum, amax);
    if (threadIdx.x == 0) band[(size_t)pair * nb + cb] = make_double2(asum, (double)amax);
    __syncthreads();
    // T[a][c0 + n] for the row tiles of this wave
    const int tiles = Np / 16;
    for (int mt = wave; mt < tiles; mt += kUpBandThreads / kWave) {
        floatx4 aR = {0.0f, 0.0f, 0.0f, 0.0f}, aI = aR;
        const float *hrow = gx + (16 * mt + lm - lk + Np);
        for (int k16 = 0; k16 < Np; k16 += 16)
#pragma unroll
            for (int k0 = k16; k0 < k16 + 16; k0 += 4) {
                const float a = hrow[-k0];
                const int bi = (k0 + lk) * 16 + lm;
                aR = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xr[bi], aR, 0, 0, 0);
                if (DTYPE == 0) aI = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xi[bi], aI, 0, 0, 0);
            }
        const int c = c0 + lm;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int a = 16 * mt + 4 * lk + i;
            if (a < N && c < N) T[((size_t)pair * N + a) * N + c] = make_float2(aR[i], aI[i]);
        }
    }
}

template <int DTYPE, int QMODE>
__global__ __launch_bounds__(kUpBandThreads) void k_up2_rows(
    const PairDesc *__restrict__ desc, int role, const int2 *__restrict__ window, int N, int Np,
    const float *__restrict__ g, const float2 *__restrict__ T, const double2 *__restrict__ band,
    int nb, float *__restrict__ amp, uint16_t *__restrict__ qout, uint16_t *__restrict__ qlo,
    int cs, double2 *__restrict__ rowstats, MixedStats *__restrict__ stats)
{
    extern __shared__ float ub[];
    const int S = Np + 1;   // odd row stride: the 16 rows of an A fragment hit distinct banks
    float *xr = ub, *xi = xr + 16 * S, *tr = xi + 16 * S, *ti = tr + 16 * S, *gx = ti + 16 * S;
    __shared__ double red[kUpBandThreads / kWave], red2[kUpBandThreads / kWave];
    __shared__ float redm[kUpBandThreads / kWave];
    const int pair = blockIdx.x, rb = blockIdx.y, a0 = 16 * rb;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const PairDesc d = desc[pair];
    const void *base;
    size_t stride;
    int line, sample;
    tile_origin(d, role, window, pair, base, stride, line, sample);
    const int total = 16 * Np;
    for (int k0 = threadIdx.x; k0 < total; k0 += kUp2Batch * blockDim.x) {
        float re[kUp2Batch], im[kUp2Batch];
        float2 tv[kUp2Batch];
#pragma unroll
        for (int j = 0; j < kUp2Batch; ++j) {
            const int k = k0 + j * blockDim.x;
            const int r = k / Np, c = k - r * Np, a = a0 + r;
            re[j] = im[j] = 0.0f;
            tv[j] = make_float2(0.0f, 0.0f);
            if (k < total && a < N && c < N) {
                load_sample<DTYPE>(base, (size_t)(line + a) * stride + sample + c, re[j], im[j]);
                tv[j] = T[((size_t)pair * N + a) * N + c];
            }
        }
#pragma unroll
        for (int j = 0; j < kUp2Batch; ++j) {
            const int k = k0 + j * blockDim.x;
            if (k < total) {
                const int r = k / Np, c = k - r * Np;
                xr[r * S + c] = re[j];
                xi[r * S + c] = im[j];
                tr[r * S + c] = tv[j].x;
                ti[r * S + c] = tv[j].y;
            }
        }
    }
    for (int k = threadIdx.x; k < 2 * Np; k += blockDim.x) {
        const int t = (k - Np) % N;
        gx[k] = g[t < 0 ? t + N : t];
    }
    const int M = 2 * N;
    float qc = 0.0f, qs = 1.0f;
    if (QMODE != kUpFP32) {
        // the quantization constants of the whole coarse tile, from the column bands
        double asum = 0.0;
        float amax = 0.0f;
        for (int b = 0; b < nb; ++b) {
            const double2 v = band[(size_t)pair * nb + b];
            asum += v.x;
            amax = fmaxf(amax, (float)v.y);
        }
        qc = (float)(asum / ((double)N * N));
        qs = (float)quant_scale(QMODE, 2.0f * amax);
        if (role == 0) {
            // zero padding of this band's chip rows
            const int rows = 2 * min(16, N - a0), pad = cs - M;
            for (int k = threadIdx.x; k < rows * pad; k += blockDim.x) {
                const int p = 2 * a0 + k / pad, c = k % pad;
                const int col = c < 16 ? c : M + c;
                qout[((size_t)pair * M + p) * cs + col] = 0;
                if (qlo) qlo[((size_t)pair * M + p) * cs + col] = 0;
            }
        }
    }
    __syncthreads();
    float *out = amp ? amp + (size_t)pair * M * M : nullptr;
    double s1 = 0.0, s2 = 0.0;
    auto emit2 = [&](int p, int q, float v0, float v1) {
        if (out) *reinterpret_cast<float2 *>(out + (size_t)p * M + q) = make_float2(v0, v1);
        if (QMODE != kUpFP32) {
            const float x0 = (v0 - qc) * qs, x1 = (v1 - qc) * qs;
            const uint16_t h0 = quantize(QMODE, x0), h1 = quantize(QMODE, x1);
            uint16_t l0 = 0, l1 = 0;
            if (qlo) {
                l0 = quantize(QMODE, x0 - dequantize(QMODE, h0));
                l1 = quantize(QMODE, x1 - dequantize(QMODE, h1));
            }
            const size_t at = role == 0 ? ((size_t)pair * M + p) * cs + 16 + q
                                        : ((size_t)pair * M + p) * M + q;
            *reinterpret_cast<uint32_t *>(qout + at) = (uint32_t)h0 | ((uint32_t)h1 << 16);
            if (qlo) *reinterpret_cast<uint32_t *>(qlo + at) = (uint32_t)l0 | ((uint32_t)l1 << 16);
            if (role == 0) {
                const double b0 = ((double)dequantize(QMODE, h0) +
                                   (qlo ? (double)dequantize(QMODE, l0) : 0.0)) / qs;
                const double b1 = ((double)dequantize(QMODE, h1) +
                                   (qlo ? (double)dequantize(QMODE, l1) : 0.0)) / qs;
                s1 += b0 + b1;
                s2 += b0 * b0 + b1 * b1;
            }
        }
    };
    const int tiles = Np / 16;
    for (int bt = wave; bt < tiles; bt += kUpBandThreads / kWave) {
        const int n = 16 * bt + lm;
        const float *hrow = gx + (n - lk + Np);
        const int arow = lm * S + lk;
        floatx4 xR = {0.0f, 0.0f, 0.0f, 0.0f}, xI = xR, tR = xR, tI = xR;
        for (int k16 = 0; k16 < Np; k16 += 16)
#pragma unroll
            for (int k0 = k16; k0 < k16 + 16; k0 += 4) {
                const float b = hrow[-k0];
                const int ai = arow + k0;
                xR = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[ai], b, xR, 0, 0, 0);
                if (DTYPE == 0) xI = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[ai], b, xI, 0, 0, 0);
                tR = __builtin_amdgcn_mfma_f32_16x16x4f32(tr[ai], b, tR, 0, 0, 0);
                if (DTYPE == 0) tI = __builtin_amdgcn_mfma_f32_16x16x4f32(ti[ai], b, tI, 0, 0, 0);
            }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = 4 * lk + i, a = a0 + r;
            if (a < N && n < N) {
                const int at = r * S + n;
                emit2(2 * a, 2 * n, amplitude(xr[at], xi[at]), amplitude(xR[i], xI[i]));
                emit2(2 * a + 1, 2 * n, amplitude(tr[at], ti[at]), amplitude(tR[i], tI[i]));
            }
        }
    }
    if (QMODE != kUpFP32) {
        if (role == 0) {
            float unused = 0.0f;
            block_sum_max(red, redm, s1, unused);
            block_sum_max(red2, redm, s2, unused);
            if (threadIdx.x == 0) rowstats[(size_t)pair * gridDim.y + rb] = make_double2(s1, s2);
        } else if (rb == 0 && threadIdx.x == 0) {
            stats[pair].scale_t = qs;
            stats[pair].mean_t = qc;
        }
    }
}

// the chip statistics of a pair from its row bands, in band order, and the
// chip's quantization scale (recomputed from the column bands)
template <int QMODE>
__global__ void k_up2_finish(int pairs, int N, const double2 *__restrict__ band, int nb,
                             const double2 *__restrict__ rowstats, int nr,
                             MixedStats *__restrict__ stats)
{
    const int pair = blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= pairs) return;
    float amax = 0.0f;
    for (int b = 0; b < nb; ++b) amax = fmaxf(amax, (float)band[(size_t)pair * nb + b].y);
    double s1 = 0.0, s2 = 0.0;
    for (int b = 0; b < nr; ++b) {
        const double2 v = rowstats[(size_t)pair * nr + b];
        s1 += v.x;
        s2 += v.y;
    }
    stats[pair].sumr = s1;
    stats[pair].sumr2 = s2;
    stats[pair].scale_r = (float)quant_scale(QMODE, 2.0f * amax);
}

// ---------------------------------------------------------------------------
// mixed coarse stage in one pass: one workgroup per pair gathers and detects
// the chip into LDS (its exact mean is removed before rounding), then streams
// the window: the quantized operand with the chip's mean amplitude removed --
// any constant works there, the normalizer takes it back exactly, and the
// chip's mean is close to the window's for the same scene -- and, when
// tgt_amp is given, the FP32 amplitudes
// ---------------------------------------------------------------------------
constexpr int kGatherPrepThreads = 256;
constexpr size_t kGatherPrepLds = 64 * 1024;

// a thread's walk over the column pairs (i, 2 m), (i, 2 m + 1) of an N-row
// tile with hw column pairs per row, kGatherPrepThreads pairs apart
struct PairWalk {
    int i, m;
    int di, dm, hw;
    __device__ PairWalk(int t, int hw_) : hw(hw_)
    {
        i = t / hw;
        m = t - i * hw;
        di = kGatherPrepThreads / hw;
        dm = kGatherPrepThreads - di * hw;
    }
    __device__ void next()
    {
        i += di;
        m += dm;
        if (m >= hw) {
            m -= hw;
            ++i;
        }
    }
};

constexpr int kGatherBatch = 8;   // column pairs per thread in flight

// the amplitudes of column pairs: |z| (FP32 square root) or |x|
template <int DTYPE>
__device__ inline void load_pairs(const void *base, size_t stride, int line, int sample, int N,
                                  PairWalk &w, float2 (&z)[kGatherBatch][2])
{
#pragma unroll
    for (int q = 0; q < kGatherBatch; ++q) {
        // rows and columns clamped into the tile: in bounds, the extras unused
        const int i = min(w.i, N - 1), j0 = min(2 * w.m, N - 1), j1 = min(2 * w.m + 1, N - 1);
        const size_t row = (size_t)(line + i) * stride + sample;
        if (DTYPE == 0) {
            z[q][0] = static_cast<const float2 *>(base)[row + j0];
            z[q][1] = static_cast<const float2 *>(base)[row + j1];
        } else {
            z[q][0] = make_float2(static_cast<const float *>(base)[row + j0], 0.0f);
            z[q][1] = make_float2(static_cast<const float *>(base)[row + j1], 0.0f);
        }
        w.next();
    }
}

__device__ inline float pair_amp(float2 z) { return amplitude(z.x, z.y); }

// sum (FP64) and max over the workgroup; every thread gets the results
__device__ inline void gp_reduce(double (*red)[kGatherPrepThreads / kWave], double &a, double &b,
                                 bool maxb)
{
    for (int d = kWave / 2; d > 0; d >>= 1) {
        a += __shfl_xor(a, d, kWave);
        const double o = __shfl_xor(b, d, kWave);
        b = maxb ? fmax(b, o) : b + o;
    }
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[0][wave] = a;
        red[1][wave] = b;
    }
    __syncthreads();
    a = red[0][0];
    b = red[1][0];
    for (int k = 1; k < kGatherPrepThreads / kWave; ++k) {
        a += red[0][k];
        b = maxb ? fmax(b, red[1][k]) : b + red[1][k];
    }
}

template <int DTYPE, int QMODE>
__global__ __launch_bounds__(kGatherPrepThreads, 4) void k_gather_prep_mixed(
    const PairDesc *__restrict__ desc, int pairs, int C, int W, int CS,
    float *__restrict__ tgt_amp, void *__restrict__ chip_qv, void *__restrict__ tgt_qv,
    MixedStats *__restrict__ stats)
{
    // operand elements: two bytes (BF16, FP16) or one (FP8); pairs of them
    // are written as one word
    constexpr bool kByte = QMODE == kUpFP8;
    using Elem = std::conditional_t<kByte, uint8_t, uint16_t>;
    using Word = std::conditional_t<kByte, uint16_t, uint32_t>;
    Elem *chip_q = static_cast<Elem *>(chip_qv), *tgt_q = static_cast<Elem *>(tgt_qv);
    extern __shared__ float ra[];   // the chip's amplitudes, [C][C]
    __shared__ double red[2][kGatherPrepThreads / kWave];
    const int pair = xcd_block(blockIdx.x, pairs);
    const int t = threadIdx.x;
    const PairDesc d = desc[pair];
    const int hc = (C + 1) / 2, hwin = (W + 1) / 2;
    // the chip, and the first batch of the window (in flight meanwhile)
    PairWalk wc(t, hc), ww(t, hwin);
    const int nbc = (C * hc + kGatherBatch * kGatherPrepThreads - 1) / (kGatherBatch * kGatherPrepThreads);
    const int nbw = (W * hwin + kGatherBatch * kGatherPrepThreads - 1) / (kGatherBatch * kGatherPrepThreads);
    float2 zw[kGatherBatch][2];
    double rs = 0.0, rm = 0.0;
    for (int b = 0; b < nbc; ++b) {
        float2 zc[kGatherBatch][2];
        const int i0 = wc.i, m0 = wc.m;
        load_pairs<DTYPE>(d.ref, d.ref_stride, d.ref_line, d.ref_sample, C, wc, zc);
        if (b == 0) load_pairs<DTYPE>(d.tgt, d.tgt_stride, d.tgt_line, d.tgt_sample, W, ww, zw);
        PairWalk w = wc;
        w.i = i0;
        w.m = m0;
        float part = 0.0f, mx = 0.0f;
#pragma unroll
        for (int q = 0; q < kGatherBatch; ++q) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int j = 2 * w.m + e;
                if (w.i < C && j < C) {
                    const float a = pair_amp(zc[q][e]);
                    ra[w.i * C + j] = a;
                    part += a;
                    mx = fmaxf(mx, a);
                }
            }
            w.next();
        }
        rs += (double)part;
        rm = fmax(rm, (double)mx);
    }
    gp_reduce(red, rs, rm, true);   // also orders the LDS writes before the reads below
    const float cr = (float)(rs / ((double)C * C));
    const float ar = (float)quant_scale(QMODE, (float)rm);
    // headroom for window values far above the chip's
    const float at = (float)quant_scale(QMODE, 16.0f * (float)rm);
    auto pack = [](float lo, float hi) -> uint32_t {
        const floatx2_t x = {lo, hi};
        if constexpr (QMODE == kUpFP8)
            return pack_fp8(lo, hi);
        else if constexpr (QMODE == kUpFP16)
            return __builtin_bit_cast(uint32_t, __builtin_convertvector(x, halfx2_t));
        else
            return __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2_t));
    };
    constexpr int kBits = 8 * sizeof(Elem);
    constexpr uint32_t kLow = (1u << kBits) - 1u;
    auto dq = [](uint32_t e) -> float {
        if constexpr (QMODE == kUpFP8) return dequant_fp8(e);
        else return dequantize(QMODE, (uint16_t)e);
    };
    // the chip operand rows: 16 zeros, the C values, zeros up to CS (CS and
    // the row starts even: pairs as words); its sums in quantized units
    Word *cq = reinterpret_cast<Word *>(chip_q + (size_t)pair * C * CS);
    double s1 = 0.0, s2 = 0.0;
    {
        PairWalk w(t, CS / 2);
        for (int k = t; k < C * (CS / 2); k += kGatherPrepThreads) {
            const int j = 2 * w.m - 16;
            uint32_t h = 0;
            if (j >= 0 && j < C) {
                const float x0 = (ra[w.i * C + j] - cr) * ar;
                const float x1 = j + 1 < C ? (ra[w.i * C + j + 1] - cr) * ar : 0.0f;
                h = pack(x0, x1);
                if (j + 1 >= C) h &= kLow;
                const double b0 = (double)dq(h & kLow);
                const double b1 = (double)dq((h >> kBits) & kLow);
                s1 += b0 + b1;
                s2 += b0 * b0 + b1 * b1;
            }
            cq[k] = (Word)h;
            w.next();
        }
    }
    // the window: quantized with the chip's mean removed, two columns per
    // 32-bit word when W is even; the next batch in flight while one is written
    float *ga = tgt_amp ? tgt_amp + (size_t)pair * W * W : nullptr;
    Elem *tq = tgt_q + (size_t)pair * W * W;
    PairWalk wo(t, hwin);
    for (int b = 0; b < nbw; ++b) {
        float2 zn[kGatherBatch][2];
        if (b + 1 < nbw) load_pairs<DTYPE>(d.tgt, d.tgt_stride, d.tgt_line, d.tgt_sample, W, ww, zn);
#pragma unroll
        for (int q = 0; q < kGatherBatch; ++q) {
            const int i = wo.i, j = 2 * wo.m;
            if (i < W) {
                const float a0 = pair_amp(zw[q][0]), a1 = pair_amp(zw[q][1]);
                const uint32_t h = pack((a0 - cr) * at, (a1 - cr) * at);
                if (ga) {
                    ga[i * W + j] = a0;
                    if (j + 1 < W) ga[i * W + j + 1] = a1;
                }
                if ((W & 1) == 0) {
                    *reinterpret_cast<Word *>(tq + i * W + j) = (Word)h;
                } else {
                    tq[i * W + j] = (Elem)h;
                    if (j + 1 < W) tq[i * W + j + 1] = (Elem)(h >> kBits);
                }
            }
            wo.next();
        }
#pragma unroll
        for (int q = 0; q < kGatherBatch; ++q) {
            zw[q][0] = zn[q][0];
            zw[q][1] = zn[q][1];
        }
    }
    gp_reduce(red, s1, s2, false);
    if (t == 0) {
        MixedStats st;
        st.sumr = s1 / ar;
        st.sumr2 = s2 / ((double)ar * ar);
        st.scale_r = ar;
        st.scale_t = at;
        st.mean_t = cr;
        stats[pair] = st;
    }
}

// ---------------------------------------------------------------------------
// zoom + sub-pixel peak + offsets.  The zoom is the band-limited (periodic
// Dirichlet) interpolation of the refined surface g (R x R, NaN read as 0) on
// the Z x Z grid, Z = zf R: H[p][b] = sum_a D(p - zf a) g[a][b], f[p][q] =
// sum_b H[p][b] D(q - zf b).  The offset comes from the first maximum of f
// over the WHOLE grid (the reference zooms the whole surface and takes its
// maximum, Sequential.icc:122-124) and a three-point parabola per axis
// through its neighbours (none at the grid's edge).  One workgroup per pair:
//
//   1. exact: f in FP64 on the (2 zf + 1)^2 cells around zf times the
//      surface's argmax, in the same order as the oracle; f_w = their maximum;
//   2. screen: f on every cell with two FP16 matrix-core products
//      (v_mfma_f32_16x16x32_f16, FP32 accumulate; the Dirichlet matrix comes
//      as precomputed fragments that serve as the A operand of H = D g and the
//      B operand of f = H D^T), H kept in FP16 in LDS one chunk of rows at a
//      time.  The host bounds its error: |f_screen - f| <= ecoef max|g| + eabs;
//   3. every cell outside the window whose screened value reaches f_w minus
//      that bound is evaluated exactly (same order), so no cell the screen
//      passes over can beat the exact maximum: the result is the first
//      maximum of the exact f over the whole grid.
// ---------------------------------------------------------------------------
constexpr int kZoomThreads = 256;
constexpr int kZoomCand = 512;   // candidates listed per pair (more are evaluated in place)

// the first maximum over a workgroup of kZoomThreads: every thread gets it
__device__ inline void block_argmax(float *rb, int *ri, float &best, int &bidx)
{
    wave_argmax(best, bidx);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        rb[wave] = best;
        ri[wave] = bidx;
    }
    __syncthreads();
    best = rb[0];
    bidx = ri[0];
    for (int w = 1; w < kZoomThreads / kWave; ++w) better(best, bidx, rb[w], ri[w]);
    __syncthreads();
}

__device__ inline void better_d(double &best, int &bidx, double x, int k)
{
    if (x > best || (x == best && k < bidx)) {
        best = x;
        bidx = k;
    }
}

// the first maximum (value, then lowest index) over the workgroup, FP64
__device__ inline void block_argmax_d(double *rb, int *ri, double &best, int &bidx)
{
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const double ob = __shfl_xor(best, d, kWave);
        const int oi = __shfl_xor(bidx, d, kWave);
        better_d(best, bidx, ob, oi);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        rb[wave] = best;
        ri[wave] = bidx;
    }
    __syncthreads();
    best = rb[0];
    bidx = ri[0];
    for (int w = 1; w < kZoomThreads / kWave; ++w) better_d(best, bidx, rb[w], ri[w]);
    __syncthreads();
}

__device__ inline double parabola_vertex(double fm, double f0, double fp)
{
    const double den = fm - 2.0 * f0 + fp;
    double d = 0.0;
    if (den < 0.0) d = 0.5 * (fm - fp) / den;
    return d < -0.5 ? -0.5 : (d > 0.5 ? 0.5 : d);
}

// LDS layout of k_zoom (byte offsets, 16-byte aligned): the exact window's
// H [span][R] and f [span][span] (FP64), the Dirichlet table [2 Z + 1] when
// staged, g [R][R] (FP32), g transposed [BT16][KP + 8] and a chunk of H
// [rt 16][KP + 8] (FP16); KP = R padded to 32, BT16 = R padded to 16
struct ZoomLayout {
    size_t f, dt, g, gt, h16, total;
};
__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }
__host__ __device__ inline ZoomLayout zoom_layout(int R, int zf, bool stage_d, int rt)
{
    const size_t span = 2 * zf + 1, Z = (size_t)zf * R;
    const size_t KP8 = (R + 31) / 32 * 32 + 8, BT16 = (R + 15) / 16 * 16;
    ZoomLayout L;
    L.f = align16(sizeof(double) * span * R);
    L.dt = align16(L.f + sizeof(double) * span * span);
    L.g = align16(L.dt + (stage_d ? sizeof(double) * (2 * Z + 1) : 0));
    L.gt = align16(L.g + sizeof(float) * R * R);
    L.h16 = align16(L.gt + 2 * BT16 * KP8);
    L.total = L.h16 + 2 * (size_t)rt * 16 * KP8;
    return L;
}

// the whole search of one pair (window, screen, exact cells): for the pairs
// flagged for it (flag[pair] != 0: no batched screen for this geometry, a
// block screen's hit, or more listed cells than the list holds)
template <bool STAGE, int KC>
__global__ __launch_bounds__(kZoomThreads) void k_zoom_full(
    const float *__restrict__ gamma_refined, int R, int zf, const double *__restrict__ D,
    const uint4 *__restrict__ dfrag, double ecoef, double eabs, int rt,
    const int2 *__restrict__ window,
    const float *__restrict__ coarse_peak, int margin, int rf, float *__restrict__ offsets,
    float *__restrict__ refined_peak, int2 *__restrict__ zoom_loc, unsigned *__restrict__ stats,
    const int *__restrict__ flag)
{
    if (flag[blockIdx.x] == 0) return;
    if (stats && threadIdx.x == 0) atomicAdd(&stats[2], 1u);
    constexpr int NQ = KC <= 2 ? 5 : 2;   // column tiles whose fragments a wave holds
    extern __shared__ __align__(16) unsigned char zlds[];
    __shared__ float rb[kZoomThreads / kWave];
    __shared__ int ri[kZoomThreads / kWave];
    __shared__ double rbd[kZoomThreads / kWave];
    __shared__ int cand[kZoomCand];
    __shared__ int ncand;
    const int pair = blockIdx.x;
    const int t = threadIdx.x, lane = t & (kWave - 1), wave = t >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int Z = zf * R;
    const int span = 2 * zf + 1;
    const int KP8 = KC * 32 + 8, BT = (R + 15) / 16, ZT = (Z + 15) / 16;
    const ZoomLayout L = zoom_layout(R, zf, STAGE, rt);
    double *H = reinterpret_cast<double *>(zlds);               // [span][R]
    double *f = reinterpret_cast<double *>(zlds + L.f);         // [span][span]
    double *Dt = reinterpret_cast<double *>(zlds + L.dt);       // [2 Z + 1], when staged
    float *g = reinterpret_cast<float *>(zlds + L.g);           // [R][R]
    _Float16 *gT = reinterpret_cast<_Float16 *>(zlds + L.gt);   // [BT 16][KP8]
    _Float16 *H16 = reinterpret_cast<_Float16 *>(zlds + L.h16); // [rt 16][KP8]

    if (STAGE)
        for (int k = t; k < 2 * Z + 1; k += kZoomThreads) Dt[k] = D[k];
    const double *Dw = STAGE ? Dt : D;   // weights, LDS or global
    const float *gr = gamma_refined + (size_t)pair * R * R;
    float best = -INFINITY, gmax = 0.0f;
    int bidx = 0x7fffffff;
    for (int k = t; k < R * R; k += kZoomThreads) {
        const float x = gr[k];
        g[k] = (x != x) ? 0.0f : x;
        if (x > best) {
            best = x;
            bidx = k;
        }
        gmax = fmaxf(gmax, fabsf(g[k]));
    }
    // g transposed in FP16 (b-major), zero outside R x R
    for (int k = t; k < BT * 16 * KC * 32; k += kZoomThreads) {
        const int b = k / (KC * 32), a = k - b * (KC * 32);
        float x = (a < R && b < R) ? gr[a * R + b] : 0.0f;
        gT[b * KP8 + a] = (_Float16)((x != x) ? 0.0f : x);
    }
    // H chunk columns past the product's 16-column tiles stay zero
    for (int k = t; k < rt * 16 * (KC * 32 - BT * 16); k += kZoomThreads) {
        const int w = KC * 32 - BT * 16, r = k / w, c = BT * 16 + (k - r * w);
        H16[r * KP8 + c] = (_Float16)0.0f;
    }
    if (t == 0) ncand = 0;
    // max |g| (for the screen's error bound), then the surface's argmax
    {
        float m = gmax;
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, kWave));
        if (lane == 0) rb[wave] = m;
        __syncthreads();
        gmax = fmaxf(fmaxf(rb[0], rb[1]), fmaxf(rb[2], rb[3]));
        __syncthreads();
    }
    block_argmax(rb, ri, best, bidx);
    const bool none = (bidx == 0x7fffffff);
    const int a = none ? 0 : bidx / R;
    const int bcol = none ? 0 : bidx - a * R;

    // 1. the exact window
    const int p0 = max(zf * a - zf, 0), p1 = min(zf * a + zf, Z - 1);
    const int q0 = max(zf * bcol - zf, 0), q1 = min(zf * bcol + zf, Z - 1);
    const int np = p1 - p0 + 1, nq = q1 - q0 + 1;
    for (int k = t; k < np * R; k += kZoomThreads) {
        const int ip = k / R, bb = k - ip * R;
        const double *w = Dw + (p0 + ip) + Z;   // weight of surface row aa: w[-zf aa]
        double acc = 0.0;
#pragma unroll 8
        for (int aa = 0; aa < R; ++aa) acc += w[-zf * aa] * (double)g[aa * R + bb];
        H[k] = acc;
    }
    __syncthreads();
    double fb = -INFINITY;
    int fi = 0x7fffffff;
    for (int k = t; k < np * nq; k += kZoomThreads) {
        const int ip = k / nq, iq = k - ip * nq;
        const double *w = Dw + (q0 + iq) + Z;
        double acc = 0.0;
#pragma unroll 8
        for (int bb = 0; bb < R; ++bb) acc += H[ip * R + bb] * w[-zf * bb];
        f[k] = acc;
        better_d(fb, fi, acc, (p0 + ip) * Z + q0 + iq);
    }
    block_argmax_d(rbd, ri, fb, fi);   // (also orders the writes of f before the reads below)

    // the exact zoomed value of any cell, in the window's order
    auto exact = [&](int p, int q) -> double {
        const double *wp = Dw + p + Z, *wq = Dw + q + Z;
        double acc = 0.0;
        for (int bb = 0; bb < R; ++bb) {
            double h = 0.0;
            for (int aa = 0; aa < R; ++aa) h += wp[-zf * aa] * (double)g[aa * R + bb];
            acc += h * wq[-zf * bb];
        }
        return acc;
    };

    // 2. the screen over the whole grid, 3. exact values where it reaches
    const double thr = fb - (ecoef * (double)gmax + eabs);
    double cb = -INFINITY;
    int ci = 0x7fffffff;
    if (!none) {
        for (int pt0 = 0; pt0 < ZT; pt0 += rt) {
            const int nrt = min(rt, ZT - pt0);
            // H = D g for this chunk's rows
            for (int item = wave; item < nrt * BT; item += kZoomThreads / kWave) {
                const int i = item / BT, bt = item - i * BT;
                floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kc = 0; kc < KC; ++kc)
                    acc = mfma_f16(dfrag[((pt0 + i) * KC + kc) * kWave + lane],
                                   *reinterpret_cast<const uint4 *>(gT + (16 * bt + lm) * KP8 +
                                                                    32 * kc + 8 * lk),
                                   acc);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    H16[(16 * i + 4 * lk + r) * KP8 + 16 * bt + lm] = (_Float16)acc[r];
            }
            __syncthreads();
            // f = H D^T: a wave holds the fragments of NQ column tiles per pass
            for (int qg = wave; qg < ZT; qg += NQ * (kZoomThreads / kWave)) {
                uint4 bq[NQ][KC];
#pragma unroll
                for (int j = 0; j < NQ; ++j)
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc) {
                        const int qt = qg + j * (kZoomThreads / kWave);
                        bq[j][kc] = qt < ZT ? dfrag[(qt * KC + kc) * kWave + lane]
                                            : make_uint4(0, 0, 0, 0);
                    }
                for (int i = 0; i < nrt; ++i) {
                    uint4 af[KC];
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc)
                        af[kc] = *reinterpret_cast<const uint4 *>(H16 + (16 * i + lm) * KP8 +
                                                                  32 * kc + 8 * lk);
#pragma unroll
                    for (int j = 0; j < NQ; ++j) {
                        const int qt = qg + j * (kZoomThreads / kWave);
                        if (qt >= ZT) break;
                        floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                        for (int kc = 0; kc < KC; ++kc) acc = mfma_f16(af[kc], bq[j][kc], acc);
                        const int q = 16 * qt + lm;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int p = 16 * (pt0 + i) + 4 * lk + r;
                            if (p >= Z || q >= Z || (double)acc[r] < thr) continue;
                            if (p >= p0 && p <= p1 && q >= q0 && q <= q1) continue;
                            const int slot = atomicAdd(&ncand, 1);
                            if (slot < kZoomCand) cand[slot] = p * Z + q;
                            else better_d(cb, ci, exact(p, q), p * Z + q);
                        }
                    }
                }
            }
            __syncthreads();
        }
        const int n = min(ncand, kZoomCand);
        for (int k = t; k < n; k += kZoomThreads) {
            const int idx = cand[k];
            better_d(cb, ci, exact(idx / Z, idx - (idx / Z) * Z), idx);
        }
        if (stats && t == 0 && ncand > 0) {
            atomicAdd(&stats[0], (unsigned)ncand);
            atomicAdd(&stats[1], 1u);
        }
    }
    block_argmax_d(rbd, ri, cb, ci);
    better_d(fb, fi, cb, ci);
    if (t != 0) return;
    if (fi == 0x7fffffff) fi = 0;
    const int pp = fi / Z, pq = fi - pp * Z;
    auto value = [&](int p, int q) -> double {
        if (p >= p0 && p <= p1 && q >= q0 && q <= q1) return f[(p - p0) * nq + (q - q0)];
        return exact(p, q);
    };
    const double dp = (pp > 0 && pp < Z - 1)
                          ? parabola_vertex(value(pp - 1, pq), fb, value(pp + 1, pq)) : 0.0;
    const double dq = (pq > 0 && pq < Z - 1)
                          ? parabola_vertex(value(pp, pq - 1), fb, value(pp, pq + 1)) : 0.0;
    const int2 tl = window[pair];
    const double zoom = (double)rf * zf;
    float dl = (float)((double)(tl.x - margin) + ((double)pp + dp) / zoom);
    float ds = (float)((double)(tl.y - margin) + ((double)pq + dq) / zoom);
    float fpk = (float)fb;
    const float cp = coarse_peak[pair];
    if (none || cp != cp) {
        dl = ds = fpk = __builtin_nanf("");
    }
    offsets[2 * pair] = dl;
    offsets[2 * pair + 1] = ds;
    refined_peak[pair] = fpk;
    zoom_loc[pair] = make_int2(pp, pq);
}

// per-pair state the zoom's kernels hand on: the threshold the screen must
// reach (exact window maximum minus the screen's error bound), the window
// (p0, p1, q0, q1), whether the pair needs the full search, and the cells
// outside the window that the screen passed (count, list of p Z + q)
struct ZoomHandoff {
    double *thr;
    int4 *win;
    int *flag;
    int *ncand;
    int *cand;   // [pairs][kZoomList]
};

// the exact window alone, with provisional results (final unless the screen
// lists cells for the pair or its maximum lies on a window edge inside the
// grid, where a parabola neighbour is outside the window); flags the pair for
// the full search when `force` (no screen for this geometry)
template <bool STAGE>
__global__ __launch_bounds__(kZoomThreads) void k_zoom_window(
    const float *__restrict__ gamma_refined, int R, int zf, const double *__restrict__ D,
    double ecoef, double eabs, int force, const int2 *__restrict__ window,
    const float *__restrict__ coarse_peak, int margin, int rf, float *__restrict__ offsets,
    float *__restrict__ refined_peak, int2 *__restrict__ zoom_loc, ZoomHandoff ho)
{
    extern __shared__ __align__(16) unsigned char zlds[];
    __shared__ float rb[kZoomThreads / kWave];
    __shared__ int ri[kZoomThreads / kWave];
    __shared__ double rbd[kZoomThreads / kWave];
    const int pair = blockIdx.x;
    const int t = threadIdx.x, lane = t & (kWave - 1), wave = t >> 6;
    const int Z = zf * R;
    const ZoomLayout L = zoom_layout(R, zf, STAGE, 0);
    double *H = reinterpret_cast<double *>(zlds);          // [span][R]
    double *f = reinterpret_cast<double *>(zlds + L.f);    // [span][span]
    double *Dt = reinterpret_cast<double *>(zlds + L.dt);  // [2 Z + 1], when staged
    float *g = reinterpret_cast<float *>(zlds + L.g);      // [R][R]

    if (STAGE)
        for (int k = t; k < 2 * Z + 1; k += kZoomThreads) Dt[k] = D[k];
    const double *Dw = STAGE ? Dt : D;
    const float *gr = gamma_refined + (size_t)pair * R * R;
    float best = -INFINITY, gmax = 0.0f;
    int bidx = 0x7fffffff;
    for (int k = t; k < R * R; k += kZoomThreads) {
        const float x = gr[k];
        g[k] = (x != x) ? 0.0f : x;
        if (x > best) {
            best = x;
            bidx = k;
        }
        gmax = fmaxf(gmax, fabsf(g[k]));
    }
    {
        float m = gmax;
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, kWave));
        if (lane == 0) rb[wave] = m;
        __syncthreads();
        gmax = fmaxf(fmaxf(rb[0], rb[1]), fmaxf(rb[2], rb[3]));
        __syncthreads();
    }
    block_argmax(rb, ri, best, bidx);
    const bool none = (bidx == 0x7fffffff);
    const int a = none ? 0 : bidx / R;
    const int bcol = none ? 0 : bidx - a * R;
    const int p0 = max(zf * a - zf, 0), p1 = min(zf * a + zf, Z - 1);
    const int q0 = max(zf * bcol - zf, 0), q1 = min(zf * bcol + zf, Z - 1);
    const int np = p1 - p0 + 1, nq = q1 - q0 + 1;
    for (int k = t; k < np * R; k += kZoomThreads) {
        const int ip = k / R, bb = k - ip * R;
        const double *w = Dw + (p0 + ip) + Z;   // weight of surface row aa: w[-zf aa]
        double acc = 0.0;
#pragma unroll 8
        for (int aa = 0; aa < R; ++aa) acc += w[-zf * aa] * (double)g[aa * R + bb];
        H[k] = acc;
    }
    __syncthreads();
    double fb = -INFINITY;
    int fi = 0x7fffffff;
    for (int k = t; k < np * nq; k += kZoomThreads) {
        const int ip = k / nq, iq = k - ip * nq;
        const double *w = Dw + (q0 + iq) + Z;
        double acc = 0.0;
#pragma unroll 8
        for (int bb = 0; bb < R; ++bb) acc += H[ip * R + bb] * w[-zf * bb];
        f[k] = acc;
        better_d(fb, fi, acc, (p0 + ip) * Z + q0 + iq);
    }
    block_argmax_d(rbd, ri, fb, fi);
    if (t != 0) return;
    if (fi == 0x7fffffff) fi = 0;
    const int pp = fi / Z, pq = fi - pp * Z;
    const int ip = pp - p0, iq = pq - q0;
    const double dp = (ip > 0 && ip < np - 1)
                          ? parabola_vertex(f[(ip - 1) * nq + iq], fb, f[(ip + 1) * nq + iq]) : 0.0;
    const double dq = (iq > 0 && iq < nq - 1)
                          ? parabola_vertex(f[ip * nq + iq - 1], fb, f[ip * nq + iq + 1]) : 0.0;
    const int2 tl = window[pair];
    const double zoom = (double)rf * zf;
    float dl = (float)((double)(tl.x - margin) + ((double)pp + dp) / zoom);
    float ds = (float)((double)(tl.y - margin) + ((double)pq + dq) / zoom);
    float fpk = (float)fb;
    const float cp = coarse_peak[pair];
    if (none || cp != cp) {
        dl = ds = fpk = __builtin_nanf("");
    }
    offsets[2 * pair] = dl;
    offsets[2 * pair + 1] = ds;
    refined_peak[pair] = fpk;
    zoom_loc[pair] = make_int2(pp, pq);
    ho.thr[pair] = none ? __builtin_nan("") : fb - (ecoef * (double)gmax + eabs);
    ho.win[pair] = make_int4(p0, p1, q0, q1);
    ho.flag[pair] = !none && force;
    ho.ncand[pair] = 0;
}

// the FP16 screen of the whole zoomed grid, kZoomScreenPairs pairs per
// workgroup: a wave owns the 16-row (and 16-column) tiles w, w + 4, ... and
// keeps their Dirichlet fragments in registers for every pair, since they are
// the A operand of H = D g for its row tiles and the B operand of f = H D^T
// for its column tiles; H goes through LDS in FP16.  A pair any of whose
// cells outside the window reaches the threshold is flagged for the full search.
constexpr int kZoomScreenPairs = 8;

__host__ __device__ inline size_t zoom_screen_lds(int R, int zf)
{
    const size_t KP8 = (R + 31) / 32 * 32 + 8;
    return 2 * KP8 * ((size_t)(R + 15) / 16 * 16 + (size_t)(zf * R + 15) / 16 * 16);
}

constexpr int kZoomScreenLoads = 5;   // surface values per thread: R^2 <= 5 x 256

template <int KC, int NT>
__global__ __launch_bounds__(kZoomThreads, 1) void k_zoom_screen(
    const float *__restrict__ gamma_refined, int pairs, int R, int zf,
    const uint4 *__restrict__ dfrag, ZoomHandoff ho)
{
    extern __shared__ __align__(16) unsigned char zlds[];
    __shared__ int hit;
    __shared__ double sthr[kZoomScreenPairs];
    __shared__ int4 swin[kZoomScreenPairs];
    __shared__ int sgo[kZoomScreenPairs];
    const int t = threadIdx.x, lane = t & (kWave - 1);
    // wave-uniform, so the tile tests below are scalar branches
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lm = lane & 15, lk = lane >> 4;
    const int Z = zf * R, ZT = (Z + 15) / 16, BT = (R + 15) / 16;
    const int KP8 = KC * 32 + 8;
    const int first = blockIdx.x * kZoomScreenPairs;
    _Float16 *gT = reinterpret_cast<_Float16 *>(zlds);    // [BT 16][KP8]
    _Float16 *H16 = gT + BT * 16 * KP8;                    // [ZT 16][KP8]
    if (t < kZoomScreenPairs) {
        const int pair = first + t;
        const bool in = pair < pairs;
        const double thr = in ? ho.thr[pair] : 0.0;
        sgo[t] = in && ho.flag[pair] == 0 && thr == thr;
        sthr[t] = thr;
        swin[t] = in ? ho.win[pair] : make_int4(0, 0, 0, 0);
    }
    // zero: g's padding (never written below) and H's columns past the tiles
    for (int k = t; k < BT * 16 * KP8; k += kZoomThreads) gT[k] = (_Float16)0.0f;
    for (int k = t; k < ZT * 16 * (KC * 32 - BT * 16); k += kZoomThreads) {
        const int w = KC * 32 - BT * 16, r = k / w, c = BT * 16 + (k - r * w);
        H16[r * KP8 + c] = (_Float16)0.0f;
    }
    uint4 fr[NT][KC];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
            const int tile = wave + 4 * j;
            fr[j][kc] = tile < ZT ? dfrag[(tile * KC + kc) * kWave + lane] : make_uint4(0, 0, 0, 0);
        }
    __syncthreads();
    // the surfaces go through registers, the next pair's loading during this one's products
    float x[kZoomScreenLoads];
    auto fetch = [&](int k) {
        const float *gr = gamma_refined + (size_t)(first + k) * R * R;
#pragma unroll
        for (int j = 0; j < kZoomScreenLoads; ++j) {
            const int i = t + j * kZoomThreads;
            x[j] = i < R * R ? gr[i] : 0.0f;
        }
    };
    int k = 0;
    while (k < kZoomScreenPairs && !sgo[k]) ++k;
    if (k < kZoomScreenPairs) fetch(k);
    while (k < kZoomScreenPairs) {
        const double thr = sthr[k];
        const int4 w = swin[k];
        const int pair = first + k;
#pragma unroll
        for (int j = 0; j < kZoomScreenLoads; ++j) {
            const int i = t + j * kZoomThreads;
            if (i < R * R) {
                const int a = i / R, b = i - a * R;
                gT[b * KP8 + a] = (_Float16)((x[j] != x[j]) ? 0.0f : x[j]);
            }
        }
        int next = k + 1;
        while (next < kZoomScreenPairs && !sgo[next]) ++next;
        if (t == 0) hit = 0;
        __syncthreads();
        // H = D g, this wave's row tiles
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            const int pt = wave + 4 * j;
            if (pt < ZT)
                for (int bt = 0; bt < BT; ++bt) {
                    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc)
                        acc = mfma_f16(fr[j][kc],
                                       *reinterpret_cast<const uint4 *>(gT + (16 * bt + lm) * KP8 +
                                                                        32 * kc + 8 * lk),
                                       acc);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        H16[(16 * pt + 4 * lk + r) * KP8 + 16 * bt + lm] = (_Float16)acc[r];
                }
        }
        __syncthreads();
        // f = H D^T: every row tile, this wave's column tiles.  The threshold is
        // rounded down to FP32 (a superset of the cells it passes); tiles inside
        // the grid and clear of the window only raise a running maximum, the
        // others are checked cell by cell
        const float thr_f = (float)(thr - fabs(thr) * 0x1p-22 - 1e-30);
        float mx = -INFINITY;
        bool any = false;
        for (int pt = 0; pt < ZT; ++pt) {
            uint4 af[KC];
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
                af[kc] = *reinterpret_cast<const uint4 *>(H16 + (16 * pt + lm) * KP8 + 32 * kc +
                                                          8 * lk);
            // every product of this row tile first (tiles past the grid have zero
            // fragments), then their results
            floatx4 acc[NT];
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                acc[j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kc = 0; kc < KC; ++kc) acc[j] = mfma_f16(af[kc], fr[j][kc], acc[j]);
            }
            const bool row_edge = 16 * pt + 15 >= Z;
            const bool row_win = 16 * pt <= w.y && 16 * pt + 15 >= w.x;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const int qt = wave + 4 * j;
                if (qt >= ZT) continue;
                const bool slow = row_edge || 16 * qt + 15 >= Z ||
                                  (row_win && 16 * qt <= w.w && 16 * qt + 15 >= w.z);
                if (!slow) {
                    mx = fmaxf(mx, fmaxf(fmaxf(acc[j][0], acc[j][1]), fmaxf(acc[j][2], acc[j][3])));
                    continue;
                }
                const int q = 16 * qt + lm;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int p = 16 * pt + 4 * lk + r;
                    const bool inside = p >= w.x && p <= w.y && q >= w.z && q <= w.w;
                    any |= p < Z && q < Z && !inside && acc[j][r] >= thr_f;
                }
            }
        }
        // the next pair's surface loads while the block settles this one
        if (next < kZoomScreenPairs) fetch(next);
        if (any || mx >= thr_f) hit = 1;
        __syncthreads();
        if (t == 0 && hit) ho.flag[pair] = 1;
        k = next;
    }
}

// The screen with one wave per pair: f^T = D_q (H^T), H^T = g^T D_p^T.  The
// output of the first product lands in the lanes as the B operand of the
// second when the second product's K slots of lane group lk hold
// b = 32 kc + 4 lk + (0..3) and 32 kc + 16 + 4 lk + (0..3); its A operand
// (D_q) comes as fragments in that order (pfrag).  Both Dirichlet fragment
// tables sit in LDS, shared by the workgroup's waves; a wave builds its pair's
// g^T fragments from global memory and needs no barrier.  Rows whose bound
// (below) stays under the threshold skip the second product.
constexpr int kScreenWaves = 8;
constexpr int kScreenMaxTiles = 40;   // row tiles: Z <= 640

__host__ __device__ inline size_t zoom_screen1_lds(int R, int zf)
{
    const size_t KC = (R + 31) / 32, ZT = ((size_t)zf * R + 15) / 16;
    return 2 * ZT * KC * kWave * 16;
}

template <int KC>
__global__ __launch_bounds__(kScreenWaves * kWave, 4) void k_zoom_screen1(
    const float *__restrict__ gamma_refined, int pairs, int R, int zf,
    const uint4 *__restrict__ dfrag, const uint4 *__restrict__ pfrag, float lambda, float hcoef,
    float habs, ZoomHandoff ho)
{
    extern __shared__ __align__(16) unsigned char zlds[];
    const int t = threadIdx.x, lane = t & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int lm = lane & 15, lk = lane >> 4;
    const int Z = zf * R, ZT = (Z + 15) / 16;
    uint4 *Dp = reinterpret_cast<uint4 *>(zlds);   // [ZT][KC][64]
    uint4 *Dq = Dp + ZT * KC * kWave;               // [ZT][KC][64]
    for (int k = t; k < ZT * KC * kWave; k += kScreenWaves * kWave) {
        Dp[k] = dfrag[k];
        Dq[k] = pfrag[k];
    }
    __syncthreads();
    const int waves = gridDim.x * kScreenWaves;
    for (int pair = blockIdx.x * kScreenWaves + wave; pair < pairs; pair += waves) {
        if (ho.flag[pair] != 0) continue;
        const double thr = ho.thr[pair];
        if (thr != thr) continue;
        const int4 w = ho.win[pair];
        const float thr_f = (float)(thr - fabs(thr) * 0x1p-22 - 1e-30);
        // column tiles: past the grid (or its last, partial one), or meeting the window
        const uint64_t valid_cols = ZT >= 64 ? ~0ull : (1ull << ZT) - 1;
        uint64_t edge_cols = ~valid_cols, win_cols = 0;
        if (Z % 16) edge_cols |= 1ull << (ZT - 1);
        for (int j = w.z / 16; j <= w.w / 16; ++j) win_cols |= 1ull << j;
        // the first product's A operand: g^T rows b = 16 bt + lm, K = a
        const float *gr = gamma_refined + (size_t)pair * R * R;
        float gv[2 * KC][KC][8];
#pragma unroll
        for (int bt = 0; bt < 2 * KC; ++bt)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int a = 32 * kc + 8 * lk + e, b = 16 * bt + lm;
                    gv[bt][kc][e] = (a < R && b < R) ? gr[a * R + b] : 0.0f;
                }
        float gmax = 0.0f;
        uint4 ga[2 * KC][KC];
#pragma unroll
        for (int bt = 0; bt < 2 * KC; ++bt)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                halfx8_t v;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float x = gv[bt][kc][e] != gv[bt][kc][e] ? 0.0f : gv[bt][kc][e];
                    gmax = fmaxf(gmax, fabsf(x));
                    v[e] = (_Float16)x;
                }
                ga[bt][kc] = __builtin_bit_cast(uint4, v);
            }
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) gmax = fmaxf(gmax, __shfl_xor(gmax, d, kWave));
        // bound on |H - H_screen| (the first product's error)
        const float eh = hcoef * gmax + habs;
        // a cell outside the window that reaches the threshold goes on the
        // pair's list; past the list's end the pair gets the full search
        auto list = [&](int p, int q) {
            const int slot = atomicAdd(&ho.ncand[pair], 1);
            if (slot < kZoomList) ho.cand[(size_t)pair * kZoomList + slot] = p * Z + q;
            else ho.flag[pair] = 1;
        };
        for (int pt = 0; pt < ZT; ++pt) {
            // H^T tiles (b rows, p columns) for this row tile of the grid
            floatx4 h[2 * KC];
#pragma unroll
            for (int bt = 0; bt < 2 * KC; ++bt) {
                h[bt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kc = 0; kc < KC; ++kc)
                    h[bt] = mfma_f16(ga[bt][kc], Dp[(pt * KC + kc) * kWave + lane], h[bt]);
            }
            // row bound: for row p, f(p, q) - c = sum_b (H(p, b) - c) D(q - zf b)
            // (a row's weights sum to 1), so f <= c + lambda r with c, r the
            // midrange and half-range of H(p, .); with H's error eh added, rows
            // whose bound stays under the threshold need no second product
            float hmax = -INFINITY, hmin = INFINITY;
#pragma unroll
            for (int bt = 0; bt < 2 * KC; ++bt)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (16 * bt + 4 * lk + r < R) {
                        hmax = fmaxf(hmax, h[bt][r]);
                        hmin = fminf(hmin, h[bt][r]);
                    }
            hmax = fmaxf(hmax, __shfl_xor(hmax, 16, kWave));
            hmin = fminf(hmin, __shfl_xor(hmin, 16, kWave));
            hmax = fmaxf(hmax, __shfl_xor(hmax, 32, kWave));
            hmin = fminf(hmin, __shfl_xor(hmin, 32, kWave));
            const float c = 0.5f * (hmax + hmin), rad = 0.5f * (hmax - hmin);
            // (rounded up past the few FP32 roundings that computed it)
            const float b0 = c + eh + lambda * (rad + eh);
            const float bound = b0 + fabsf(b0) * 0x1p-20f + 0x1p-60f;
            if (!__any(16 * pt + lm < Z && bound >= thr_f)) continue;
            // as the second product's B operand, in the K order of pfrag
            uint4 hb[KC];
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
                const halfx8_t v = {(_Float16)h[2 * kc][0], (_Float16)h[2 * kc][1],
                                    (_Float16)h[2 * kc][2], (_Float16)h[2 * kc][3],
                                    (_Float16)h[2 * kc + 1][0], (_Float16)h[2 * kc + 1][1],
                                    (_Float16)h[2 * kc + 1][2], (_Float16)h[2 * kc + 1][3]};
                hb[kc] = __builtin_bit_cast(uint4, v);
            }
            const bool row_edge = 16 * pt + 15 >= Z;
            const bool row_win = 16 * pt <= w.y && 16 * pt + 15 >= w.x;
            const uint64_t slow = row_edge ? ~0ull : (edge_cols | (row_win ? win_cols : 0ull));
            // column tiles in groups of 8: their products first, then their results
            for (int j0 = 0; j0 < ZT; j0 += 8) {
                floatx4 acc[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    acc[u] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
                    const int j = min(j0 + u, ZT - 1);
#pragma unroll
                    for (int kc = 0; kc < KC; ++kc)
                        acc[u] = mfma_f16(Dq[(j * KC + kc) * kWave + lane], hb[kc], acc[u]);
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int j = j0 + u;
                    if (j >= ZT) break;
                    // lane: f[16 pt + lm][16 j + 4 lk + r]; the tiles inside the
                    // grid and clear of the window test four cells at once
                    const float m4 =
                        fmaxf(fmaxf(acc[u][0], acc[u][1]), fmaxf(acc[u][2], acc[u][3]));
                    const bool fast = !((slow >> j) & 1);
                    if (fast && !(m4 >= thr_f)) continue;
                    const int p = 16 * pt + lm;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int q = 16 * j + 4 * lk + r;
                        const bool inside = p >= w.x && p <= w.y && q >= w.z && q <= w.w;
                        if (p < Z && q < Z && !inside && acc[u][r] >= thr_f) list(p, q);
                    }
                }
            }
        }
    }
}

// the listed cells of the pairs not flagged for the full search, and the pairs
// whose window maximum lies on a window edge inside the grid: the exact value
// of each listed cell, the first maximum of the window's and the listed
// cells' values, and the sub-pixel fit from exact neighbour values.  Every
// cell outside the window that can reach the window's maximum is on the list,
// so the result is the full search's.
//
// A cell's value is f(p, q) = sum_b H(p, b) D(q - zf b) with H(p, b) =
// sum_a D(p - zf a) g(a, b), both sums in index order -- the window's order,
// so a cell's value has the bits the window or the full search computes.  The
// rows H(p, .) of the listed cells' rows (up to nr of them) are computed once
// into LDS, one lane per (row, b); a cell in a row past those takes both sums
// itself.  A workgroup takes the pairs blockIdx.x + k gridDim.x that need the
// stage, one after the other.
__host__ __device__ inline size_t zoom_cands_lds(int R, int zf, bool stage_d, int nr)
{
    const size_t Z = (size_t)zf * R;
    return (stage_d ? align16(sizeof(double) * (2 * Z + 1)) : 0) +
           align16(sizeof(float) * (size_t)R * R) + align16(sizeof(int) * Z) +
           sizeof(double) * (size_t)nr * R;
}

constexpr int kZoomSharedRows = 64;   // at most kZoomThreads
constexpr int kZoomCandBlocks = 2048;  // workgroups of the stage (each looks up its pairs at once)

template <bool STAGE>
__global__ __launch_bounds__(kZoomThreads) void k_zoom_cands(
    const float *__restrict__ gamma_refined, int pairs, int R, int zf,
    const double *__restrict__ D, int nr, const int2 *__restrict__ window,
    const float *__restrict__ coarse_peak, int margin, int rf, float *__restrict__ offsets,
    float *__restrict__ refined_peak, int2 *__restrict__ zoom_loc, unsigned *__restrict__ stats,
    ZoomHandoff ho)
{
    extern __shared__ __align__(16) unsigned char zlds[];
    __shared__ int ri[kZoomThreads / kWave];
    __shared__ double rbd[kZoomThreads / kWave];
    __shared__ double nb[4];
    __shared__ int nrows, nact;
    __shared__ int rowlist[kZoomSharedRows];
    __shared__ int act[kZoomThreads];
    const int t = threadIdx.x;
    const int Z = zf * R;
    size_t off = 0;
    double *Dt = reinterpret_cast<double *>(zlds);   // [2 Z + 1], when staged
    if (STAGE) off = align16(sizeof(double) * (2 * (size_t)Z + 1));
    float *g = reinterpret_cast<float *>(zlds + off);   // [R][R]
    off += align16(sizeof(float) * (size_t)R * R);
    int *rowslot = reinterpret_cast<int *>(zlds + off);   // [Z]: H slot of row p, or -1
    off += align16(sizeof(int) * (size_t)Z);
    double *Hr = reinterpret_cast<double *>(zlds + off);   // [nr][R]
    if (STAGE)
        for (int k = t; k < 2 * Z + 1; k += kZoomThreads) Dt[k] = D[k];
    const double *Dw = STAGE ? Dt : D;

    // the pairs base + k gridDim.x (k < kZoomThreads) of this workgroup that
    // need the stage: looked up at once, then taken one after the other
    for (int base = blockIdx.x; base < pairs; base += gridDim.x * kZoomThreads) {
        if (t == 0) nact = 0;
        __syncthreads();
        {
            const int pair = base + t * gridDim.x;
            if (pair < pairs && ho.flag[pair] == 0) {
                const double thr = ho.thr[pair];
                const int4 w = ho.win[pair];
                const int2 loc = zoom_loc[pair];
                const bool edge = (loc.x == w.x && w.x > 0) || (loc.x == w.y && w.y < Z - 1) ||
                                  (loc.y == w.z && w.z > 0) || (loc.y == w.w && w.w < Z - 1);
                if (thr == thr && (ho.ncand[pair] > 0 || edge)) act[atomicAdd(&nact, 1)] = pair;
            }
        }
        __syncthreads();
        const int na = nact;
        for (int ia = 0; ia < na; ++ia) {
            const int pair = act[ia];
            const int2 loc = zoom_loc[pair];
            const int n = min(ho.ncand[pair], kZoomList);
            const float *gr = gamma_refined + (size_t)pair * R * R;
            for (int k = t; k < R * R; k += kZoomThreads) {
                const float x = gr[k];
                g[k] = (x != x) ? 0.0f : x;
            }
            for (int k = t; k < Z; k += kZoomThreads) rowslot[k] = -1;
            if (t == 0) nrows = 0;
            __syncthreads();
            // the rows to share: the listed cells' and the window maximum's
            const int *cl = ho.cand + (size_t)pair * kZoomList;
            for (int k = t; k < n; k += kZoomThreads) rowslot[cl[k] / Z] = -2;
            if (t == 0) rowslot[loc.x] = -2;
            __syncthreads();
            for (int p = t; p < Z; p += kZoomThreads) {
                if (rowslot[p] != -2) continue;
                const int slot = atomicAdd(&nrows, 1);
                if (slot < nr) rowlist[slot] = p;
                rowslot[p] = slot < nr ? slot : -1;
            }
            __syncthreads();
            // H of shared rows [s0, s1)
            auto rows_h = [&](int s0, int s1) {
                for (int k = t; k < (s1 - s0) * R; k += kZoomThreads) {
                    const int sl = s0 + k / R, bb = k - (sl - s0) * R;
                    const double *wp = Dw + rowlist[sl] + Z;
                    double h = 0.0;
#pragma unroll 8
                    for (int aa = 0; aa < R; ++aa) h += wp[-zf * aa] * (double)g[aa * R + bb];
                    Hr[sl * R + bb] = h;
                }
            };
            const int ns = min(nrows, nr);
            rows_h(0, ns);
            __syncthreads();
            auto value = [&](int p, int q) -> double {
                const double *wq = Dw + q + Z;
                const int sl = rowslot[p];
                double acc = 0.0;
                if (sl >= 0) {
#pragma unroll 8
                    for (int bb = 0; bb < R; ++bb) acc += Hr[sl * R + bb] * wq[-zf * bb];
                } else {
                    const double *wp = Dw + p + Z;
                    for (int bb = 0; bb < R; ++bb) {
                        double h = 0.0;
#pragma unroll 8
                        for (int aa = 0; aa < R; ++aa) h += wp[-zf * aa] * (double)g[aa * R + bb];
                        acc += h * wq[-zf * bb];
                    }
                }
                return acc;
            };
            double cb = -INFINITY;
            int ci = 0x7fffffff;
            for (int k = t; k < n; k += kZoomThreads) {
                const int idx = cl[k];
                better_d(cb, ci, value(idx / Z, idx - (idx / Z) * Z), idx);
            }
            if (stats && t == 0 && n > 0) {
                atomicAdd(&stats[0], (unsigned)n);
                atomicAdd(&stats[1], 1u);
            }
            // the window's maximum (its value recomputed)
            double fb = value(loc.x, loc.y);
            int fi = loc.x * Z + loc.y;
            block_argmax_d(rbd, ri, cb, ci);
            better_d(fb, fi, cb, ci);
            const int pp = fi / Z, pq = fi - pp * Z;
            // the rows of the neighbours above and below join the shared rows
            // while there is room
            if (t == 0) {
                nrows = ns;
                for (int r = pp - 1; r <= pp + 1; r += 2)
                    if (r >= 0 && r < Z && rowslot[r] < 0 && nrows < nr) {
                        rowslot[r] = nrows;
                        rowlist[nrows++] = r;
                    }
            }
            __syncthreads();
            rows_h(ns, nrows);
            __syncthreads();
            if (t < 4) {
                const int p = pp + (t == 0 ? -1 : t == 1 ? 1 : 0);
                const int q = pq + (t == 2 ? -1 : t == 3 ? 1 : 0);
                nb[t] = (p >= 0 && p < Z && q >= 0 && q < Z) ? value(p, q) : 0.0;
            }
            __syncthreads();
            if (t == 0) {
                const double dp = (pp > 0 && pp < Z - 1) ? parabola_vertex(nb[0], fb, nb[1]) : 0.0;
                const double dq = (pq > 0 && pq < Z - 1) ? parabola_vertex(nb[2], fb, nb[3]) : 0.0;
                const int2 tl = window[pair];
                const double zoom = (double)rf * zf;
                float dl = (float)((double)(tl.x - margin) + ((double)pp + dp) / zoom);
                float ds = (float)((double)(tl.y - margin) + ((double)pq + dq) / zoom);
                float fpk = (float)fb;
                const float cp = coarse_peak[pair];
                if (cp != cp) {
                    dl = ds = fpk = __builtin_nanf("");
                }
                offsets[2 * pair] = dl;
                offsets[2 * pair + 1] = ds;
                refined_peak[pair] = fpk;
                zoom_loc[pair] = make_int2(pp, pq);
            }
            __syncthreads();   // (the next pair reuses the LDS)
        }
        __syncthreads();   // (the next lookup reuses act and nact)
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static size_t correlate_lds(int C, int bu, int ucount, int vcount, int jblock)
{
    const int nub = (ucount + bu - 1) / bu;
    return (size_t)correlate_stride(nub * bu + C - 1) * (vcount + jblock - 1) * sizeof(float);
}

CorrelateShape choose_correlate_shape(int C, int P, bool exact, int bu, int nsplit, int jblock)
{
    CorrelateShape s{};
    s.exact = exact;
    if (bu == 1 || bu == 2 || bu == 4 || bu == 5 || bu == 8 || bu == 16)
        s.bu = bu;
    else if (!exact && C % 4 == 0)
        s.bu = 4;  // measured best for the split refined contraction
    else
        s.bu = (C % 8 == 0) ? 8 : (C % 4 == 0) ? 4 : (C % 2 == 0) ? 2 : 1;
    const size_t budget = 64 * 1024;
    // chip-column split: the exact form keeps the whole contraction in a lane
    s.nsplit = 1;
    if (!exact) {
        if (nsplit > 0 && (nsplit & (nsplit - 1)) == 0 && nsplit <= 16 && C % nsplit == 0) {
            s.nsplit = nsplit;
        } else {
            for (int ns : {8, 4, 2})
                if (C % ns == 0 && C / ns >= 16) {
                    s.nsplit = ns;
                    break;
                }
        }
    }
    // the largest square tile whose tasks (one per thread) and LDS fit
    int extent = 1;
    const int candidates[] = {P, 96, 80, 64, 48, 32, 24, 16, 8, 4, 2, 1};
    for (int e : candidates) {
        if (e > P) continue;
        extent = e;
        const int tasks = e * ((e + s.bu - 1) / s.bu);
        if (tasks <= 1024 && correlate_lds(C, s.bu, e, e, kJB) <= budget) break;
    }
    s.ub = s.vb = extent;
    s.tiles_u = s.tiles_v = (P + extent - 1) / extent;
    // stage the whole chip-column range at once when it fits
    const int jlen = C / s.nsplit;
    s.jblock = correlate_lds(C, s.bu, extent, extent, jlen) <= 80 * 1024 ? jlen : kJB;
    // a smaller block: more workgroups per CU, more barriers
    if (jblock > 0 && jblock < s.jblock) s.jblock = jblock;
    s.lds_bytes = correlate_lds(C, s.bu, extent, extent, s.jblock);
    const int tasks = extent * ((extent + s.bu - 1) / s.bu);
    s.threads = ((tasks + kWave - 1) / kWave) * kWave;
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
                             double2 *stats, bool exact)
{
    // staged in LDS where the chip fits and its rows start 16-byte aligned:
    // up to 32 KiB in 256-thread workgroups (measured: 64-sample chips 0.100 ->
    // 0.064 ms); where the order need not be the twin's, up to 64 KiB in
    // 1024-thread workgroups (two per CU still fill it), and larger chips
    // unstaged in 1024-thread workgroups
    const size_t bytes = (size_t)C * C * sizeof(float);
    if (bytes <= 32 * 1024 && C % 4 == 0)
        hipLaunchKernelGGL((k_chip_stats<true, kStatThreads>), dim3(pairs), dim3(kStatThreads),
                           bytes, s, amp, C, rT, stats);
    else if (!exact && bytes <= 64 * 1024 && C % 4 == 0)
        hipLaunchKernelGGL((k_chip_stats<true, 1024>), dim3(pairs), dim3(1024), bytes, s, amp, C,
                           rT, stats);
    else if (!exact)
        hipLaunchKernelGGL((k_chip_stats<false, 1024>), dim3(pairs), dim3(1024), 0, s, amp, C, rT,
                           stats);
    else
        hipLaunchKernelGGL((k_chip_stats<false, kStatThreads>), dim3(pairs), dim3(kStatThreads), 0,
                           s, amp, C, rT, stats);
    return hipGetLastError();
}

hipError_t launch_window_sums(hipStream_t s, const float *amp, int pairs, int W, int C,
                              double2 *sums)
{
    const dim3 grid(pairs), block(kWave);
    switch ((W + kWave - 1) / kWave) {
    case 1: hipLaunchKernelGGL(k_window_sums<1>, grid, block, 0, s, amp, W, C, sums); break;
    case 2: hipLaunchKernelGGL(k_window_sums<2>, grid, block, 0, s, amp, W, C, sums); break;
    case 3: hipLaunchKernelGGL(k_window_sums<3>, grid, block, 0, s, amp, W, C, sums); break;
    case 4: hipLaunchKernelGGL(k_window_sums<4>, grid, block, 0, s, amp, W, C, sums); break;
    case 5: hipLaunchKernelGGL(k_window_sums<5>, grid, block, 0, s, amp, W, C, sums); break;
    case 6: hipLaunchKernelGGL(k_window_sums<6>, grid, block, 0, s, amp, W, C, sums); break;
    case 7: hipLaunchKernelGGL(k_window_sums<7>, grid, block, 0, s, amp, W, C, sums); break;
    case 8: hipLaunchKernelGGL(k_window_sums<8>, grid, block, 0, s, amp, W, C, sums); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_correlate(hipStream_t s, const CorrelateShape &sh, const float *rT,
                            const float *amp, int pairs, int C, int W, float *partial,
                            const double2 *sums, const double2 *stats, float *gamma)
{
    const int P = W - C + 1;
    const size_t blocks = (size_t)sh.tiles_u * sh.tiles_v * sh.nsplit;
    if (sh.exact && (sh.nsplit != 1 || !sums || !stats || !gamma)) return hipErrorInvalidValue;
    if (!sh.exact && !partial) return hipErrorInvalidValue;
    // a launch's work-items fit in 32 bits: pairs go in chunks
    const int chunk = std::min(pairs, launch_pair_cap(blocks * sh.threads));
    for (int first = 0; first < pairs; first += chunk) {
        const int n = std::min(chunk, pairs - first);
        const dim3 grid((unsigned)(n * blocks)), block(sh.threads);
        const float *rTf = rT + (size_t)first * C * C;
        const float *ampf = amp + (size_t)first * W * W;
        float *pf = sh.exact ? nullptr : partial + (size_t)first * sh.nsplit * P * P;
        const double2 *sf = sh.exact ? sums + (size_t)first * P * P : nullptr;
        const double2 *stf = sh.exact ? stats + first : nullptr;
        float *gf = sh.exact ? gamma + (size_t)first * P * P : nullptr;
#define AMPCOR_CORRELATE(BUV, EX)                                                                \
    hipLaunchKernelGGL((k_correlate<BUV, EX>), grid, block, sh.lds_bytes, s, rTf, ampf, C, W, P, \
                       sh.ub, sh.vb, sh.tiles_u, sh.tiles_v, sh.nsplit, sh.jblock, pf, sf, stf, gf)
#define AMPCOR_CORRELATE_BU(EX)                      \
    switch (sh.bu) {                                 \
    case 16: AMPCOR_CORRELATE(16, EX); break;        \
    case 8: AMPCOR_CORRELATE(8, EX); break;          \
    case 5: AMPCOR_CORRELATE(5, EX); break;          \
    case 4: AMPCOR_CORRELATE(4, EX); break;          \
    case 2: AMPCOR_CORRELATE(2, EX); break;          \
    case 1: AMPCOR_CORRELATE(1, EX); break;          \
    default: return hipErrorInvalidValue;            \
    }
        if (sh.exact) {
            AMPCOR_CORRELATE_BU(true)
        } else {
            AMPCOR_CORRELATE_BU(false)
        }
#undef AMPCOR_CORRELATE_BU
#undef AMPCOR_CORRELATE
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_normalize(hipStream_t s, const float *partial, int nsplit, const double2 *sums,
                            const double2 *stats, int pairs, int C, int W, float *gamma)
{
    const int P = W - C + 1;
    const size_t pp = (size_t)P * P;
    const int chunk = std::min(pairs, launch_pair_cap(pp));
    for (int first = 0; first < pairs; first += chunk) {
        const int n = std::min(chunk, pairs - first);
        const unsigned blocks = (unsigned)((n * pp + 255) / 256);
        hipLaunchKernelGGL(k_normalize, dim3(blocks), dim3(256), 0, s,
                           partial + (size_t)first * nsplit * pp, nsplit, sums + first * pp,
                           stats + first, C, P, n, gamma + first * pp);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

bool normalize_fuses(int W, int C)
{
    return W <= kFuseThreads * kFuseMaxCols && W - C + 1 <= 16 * kFuseOut;
}

hipError_t launch_normalize_fused(hipStream_t s, int form, const void *target, int pairs, int W,
                                  int C, const float *partial, int nsplit, const double2 *stats,
                                  const MixedStats *mstats, float *gamma)
{
    if (!normalize_fuses(W, C) || (form == kFuseFP32 ? !stats : !mstats))
        return hipErrorInvalidValue;
    const bool one = W <= kFuseThreads;   // one target column per thread
    // outputs per thread and block: 3, 5 or 9 (P <= 48, 80, 144)
    const int P = W - C + 1, jo = P <= 48 ? 3 : P <= 80 ? 5 : 9;
    // wide windows with few outputs per thread: the lazy offsets (measured faster
    // for the 160-sample refined window; slower where four workgroups fit a CU)
    const bool lazy = one && jo == 3 && W >= 144;
    // a lone last placement row (P = 16 n + 1) joins the block before it in
    // the lazy-offset kernels, where the LDS of one more row leaves as many
    // workgroups per CU (their register bound: three)
    auto lds_of = [&](int rows) { return 2 * (size_t)rows * fuse_row_stride(W) * sizeof(double); };
    auto per_cu = [&](int rows) {
        const size_t fixed = lazy ? 2 * (size_t)(kFuseRows + 1) * 16 * sizeof(double) : 0;
        return (int)(kLdsPerCu / (lds_of(rows) + fixed));
    };
    const int cap = one && jo < 9 ? (lazy ? 3 : 4) : 2;
    const int xr = lazy && P > kFuseRows && (P - 1) % kFuseRows == 0 &&
                   per_cu(kFuseRows + 1) >= std::min(cap, per_cu(kFuseRows));
    const size_t lds = lds_of(kFuseRows + xr);
#define AMPCOR_FUSE_K(F, CO, J, LZ)                                                                 \
    hipLaunchKernelGGL((k_normalize_fused<F, CO, J, LZ>), dim3(pairs), dim3(kFuseThreads), lds, s,  \
                       target, W, C, partial, nsplit, stats, mstats, gamma, xr)
#define AMPCOR_FUSE_J(F, CO, J)                                                                     \
    do {                                                                                            \
        if (CO == 1 && J == 3 && lazy) AMPCOR_FUSE_K(F, CO, J, (CO == 1 && J == 3));                \
        else AMPCOR_FUSE_K(F, CO, J, false);                                                        \
    } while (0)
#define AMPCOR_FUSE_C(F, CO)                                                                        \
    do {                                                                                            \
        if (jo == 3) AMPCOR_FUSE_J(F, CO, 3);                                                       \
        else if (jo == 5) AMPCOR_FUSE_J(F, CO, 5);                                                  \
        else AMPCOR_FUSE_J(F, CO, 9);                                                               \
    } while (0)
#define AMPCOR_FUSE(F)                                                                              \
    do {                                                                                            \
        if (one) AMPCOR_FUSE_C(F, 1);                                                               \
        else AMPCOR_FUSE_C(F, 2);                                                                   \
    } while (0)
    switch (form) {
    case kFuseFP32: AMPCOR_FUSE(kFuseFP32); break;
    case kFuseMixed: AMPCOR_FUSE(kFuseMixed); break;
    case kFuseBF16: AMPCOR_FUSE(kFuseBF16); break;
    case kFuseFP16: AMPCOR_FUSE(kFuseFP16); break;
    case kFuseFP8: AMPCOR_FUSE(kFuseFP8); break;
    default: return hipErrorInvalidValue;
    }
#undef AMPCOR_FUSE
#undef AMPCOR_FUSE_C
#undef AMPCOR_FUSE_J
#undef AMPCOR_FUSE_K
    return hipGetLastError();
}

hipError_t launch_coarse_peak(hipStream_t s, const float *gamma, int pairs, int P, int W, int E,
                              int refine_margin, int exclude, int2 *loc, int2 *window,
                              float *peak, float *snr, float tau, int *near)
{
    // surfaces up to 67 x 67 stay in one wave's registers, up to 135 x 135 in
    // one workgroup's
    constexpr int kKeep = 72;
    if (P * P <= kKeep * kWave)
        hipLaunchKernelGGL(k_coarse_peak<kKeep>, dim3((pairs + 3) / 4), dim3(256), 0, s, gamma,
                           pairs, P, W, E, refine_margin, exclude, loc, window, peak, snr, tau,
                           near);
    else if (P * P <= kKeep * 256)   // one workgroup a pair, the surface in registers
        hipLaunchKernelGGL(k_coarse_peak_wide<kKeep>, dim3(pairs), dim3(256), 0, s, gamma, P, W,
                           E, refine_margin, exclude, loc, window, peak, snr, tau, near);
    else
        hipLaunchKernelGGL(k_coarse_peak_wide<0>, dim3(pairs), dim3(256), 0, s, gamma, P, W, E,
                           refine_margin, exclude, loc, window, peak, snr, tau, near);
    return hipGetLastError();
}

float peak_guard_margin(int C) { return 0.75f / (float)C; }

hipError_t launch_peak_guard(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int C,
                             int W, int E, int rm, int exclude, const float *gamma, const int *near,
                             int2 *loc, int2 *window, float *peak, float *snr, unsigned *counters)
{
    hipLaunchKernelGGL(k_guard_clear, dim3(1), dim3(64), 0, s, counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || pairs == 0) return e;
    hipLaunchKernelGGL(k_peak_guard, dim3(pairs), dim3(kGuardThreads), 0, s, desc, dtype, C, W, E,
                       rm, exclude, peak_guard_margin(C), gamma, near, loc, window, peak, snr,
                       counters);
    return hipGetLastError();
}

bool upsample_fuses_quantization(int N, int rf) { return rf == 2 && N <= kUpBandMaxN; }

// the FP32 and split forms take the banded two-pass form above 64 samples
// (measured faster: 80-sample tiles 0.80 -> 0.67 ms FP32, 0.97 -> 0.88 ms
// split, than the LDS-resident kernel, which holds one workgroup per CU there)
constexpr int kUp2ResidentMaxN = 64;

bool upsample_needs_scratch(int N, int rf)
{
    return rf == 2 && (N + 15) / 16 * 16 > kUp2ResidentMaxN && N <= kUpBandMaxN;
}

bool upsample_uses_fragments(int N) { return (N + 31) / 32 * 32 <= kUp2hMaxNReal; }

int upsample_fragment_count(int N) { return ((N + 31) / 32 * 32 / 16) * ((N + 31) / 32) * 64 * 8; }

hipError_t launch_upsample(hipStream_t s, const PairDesc *desc, int pairs, int dtype, int role,
                           const int2 *window, int N, int rf, const float *U, const float *g,
                           float *amp, int qmode, uint16_t *q, int cs, MixedStats *stats,
                           uint16_t *qlo, const uint16_t *hfrag, const UpsampleScratch *scratch)
{
    const int Np = (N + 15) / 16 * 16;
    const int Np32 = (N + 31) / 32 * 32;
    if (rf == 2 && hfrag && qmode != kUpFP32 && !qlo && Np32 <= up2h_max_n(dtype)) {
        const size_t lds = up2h_lds_bytes(Np32, dtype);
        const dim3 grid(pairs), block(kUp2Waves * kWave);
        const uint4 *hf = reinterpret_cast<const uint4 *>(hfrag);
#define AMPCOR_UP2H(DT, QM, MN)                                                                    \
    hipLaunchKernelGGL((k_upsample2h<DT, QM, MN>), grid, block, lds, s, desc, role, window, N, Np32, \
                       hf, amp, q, cs, stats)
        if (dtype == 0) {
            if (Np32 <= 64) {   // 76 VGPRs: three workgroups per CU instead of two
                if (qmode == kUpFP16) AMPCOR_UP2H(0, kUpFP16, 64);
                else AMPCOR_UP2H(0, kUpBF16, 64);
            } else {
                if (qmode == kUpFP16) AMPCOR_UP2H(0, kUpFP16, kUp2hMaxN);
                else AMPCOR_UP2H(0, kUpBF16, kUp2hMaxN);
            }
        } else {
            if (Np32 <= 128) {   // 102 VGPRs: two workgroups per CU instead of one
                if (qmode == kUpFP16) AMPCOR_UP2H(1, kUpFP16, 128);
                else AMPCOR_UP2H(1, kUpBF16, 128);
            } else {
                if (qmode == kUpFP16) AMPCOR_UP2H(1, kUpFP16, kUp2hMaxNReal);
                else AMPCOR_UP2H(1, kUpBF16, kUp2hMaxNReal);
            }
        }
#undef AMPCOR_UP2H
        return hipGetLastError();
    }
    if (upsample_needs_scratch(N, rf) && g && scratch && scratch->T) {
        const int Np = (N + 15) / 16 * 16, nb = Np / 16;
        const dim3 grid(pairs, nb), block(kUpBandThreads);
        const size_t lds1 = (2 * (size_t)Np * 16 + 2 * Np) * sizeof(float);
        const size_t lds2 = (4 * (size_t)16 * (Np + 1) + 2 * Np) * sizeof(float);
        if (dtype == 0)
            hipLaunchKernelGGL(k_up2_cols<0>, grid, block, lds1, s, desc, role, window, N, Np, g,
                               scratch->T, scratch->band);
        else
            hipLaunchKernelGGL(k_up2_cols<1>, grid, block, lds1, s, desc, role, window, N, Np, g,
                               scratch->T, scratch->band);
#define AMPCOR_ROWS(DT, QM)                                                                        \
    hipLaunchKernelGGL((k_up2_rows<DT, QM>), grid, block, lds2, s, desc, role, window, N, Np, g,    \
                       scratch->T, scratch->band, nb, amp, q, qlo, cs, scratch->rows, stats)
        if (dtype == 0) {
            if (qmode == kUpBF16) AMPCOR_ROWS(0, kUpBF16);
            else if (qmode == kUpFP16) AMPCOR_ROWS(0, kUpFP16);
            else AMPCOR_ROWS(0, kUpFP32);
        } else {
            if (qmode == kUpBF16) AMPCOR_ROWS(1, kUpBF16);
            else if (qmode == kUpFP16) AMPCOR_ROWS(1, kUpFP16);
            else AMPCOR_ROWS(1, kUpFP32);
        }
#undef AMPCOR_ROWS
        if (role == 0 && qmode != kUpFP32) {
            const dim3 fg((pairs + 255) / 256), fb(256);
            if (qmode == kUpFP16)
                hipLaunchKernelGGL(k_up2_finish<kUpFP16>, fg, fb, 0, s, pairs, N, scratch->band, nb,
                                   scratch->rows, nb, stats);
            else
                hipLaunchKernelGGL(k_up2_finish<kUpBF16>, fg, fb, 0, s, pairs, N, scratch->band, nb,
                                   scratch->rows, nb, stats);
        }
        return hipGetLastError();
    }
    if (rf == 2 && Np <= kUp2MaxN && g) {
        const size_t lds = (4 * (size_t)Np * (Np + 4) + 2 * Np) * sizeof(float);
        const dim3 grid(pairs), block(kUp2Waves * kWave);
#define AMPCOR_UP2(DT, QM)                                                                        \
    hipLaunchKernelGGL((k_upsample2<DT, QM>), grid, block, lds, s, desc, role, window, N, Np, g, \
                       amp, q, qlo, cs, stats)
        if (dtype == 0) {
            if (qmode == kUpBF16) AMPCOR_UP2(0, kUpBF16);
            else if (qmode == kUpFP16) AMPCOR_UP2(0, kUpFP16);
            else AMPCOR_UP2(0, kUpFP32);
        } else {
            if (qmode == kUpBF16) AMPCOR_UP2(1, kUpBF16);
            else if (qmode == kUpFP16) AMPCOR_UP2(1, kUpFP16);
            else AMPCOR_UP2(1, kUpFP32);
        }
#undef AMPCOR_UP2
        return hipGetLastError();
    }
    if (qmode != kUpFP32) return hipErrorInvalidValue;  // callers quantize separately
    const int M = rf * N;
    const int Mp = (M + 15) / 16 * 16;
    const size_t lds = (size_t)kUpWaves * 2 * 16 * Np * sizeof(float);
    if (dtype == 0)
        hipLaunchKernelGGL(k_upsample<0>, dim3(pairs), dim3(kUpWaves * kWave), lds, s, desc, role,
                           window, N, Np, M, Mp, U, amp);
    else
        hipLaunchKernelGGL(k_upsample<1>, dim3(pairs), dim3(kUpWaves * kWave), lds, s, desc, role,
                           window, N, Np, M, Mp, U, amp);
    return hipGetLastError();
}

bool gather_prep_fits(int C, int W) { return (size_t)C * C * sizeof(float) <= kGatherPrepLds; }

hipError_t launch_gather_prep_mixed(hipStream_t s, const PairDesc *desc, int pairs, int dtype,
                                    int C, int W, int CS, int qmode, float *tgt_amp,
                                    void *chip_q, void *tgt_q, MixedStats *stats)
{
    const size_t lds = (size_t)C * C * sizeof(float);
    const dim3 grid(pairs), block(kGatherPrepThreads);
#define AMPCOR_GP(DT, QM)                                                                            \
    hipLaunchKernelGGL((k_gather_prep_mixed<DT, QM>), grid, block, lds, s, desc, pairs, C, W, CS,     \
                       tgt_amp, chip_q, tgt_q, stats)
#define AMPCOR_GP_Q(DT)                                                                              \
    do {                                                                                             \
        if (qmode == kUpFP8) AMPCOR_GP(DT, kUpFP8);                                                  \
        else if (qmode == kUpFP16) AMPCOR_GP(DT, kUpFP16);                                           \
        else AMPCOR_GP(DT, kUpBF16);                                                                 \
    } while (0)
    if (qmode != kUpBF16 && qmode != kUpFP16 && qmode != kUpFP8) return hipErrorInvalidValue;
    if (dtype == 0) AMPCOR_GP_Q(0);
    else AMPCOR_GP_Q(1);
#undef AMPCOR_GP_Q
#undef AMPCOR_GP
    return hipGetLastError();
}

constexpr size_t kZoomLdsStage = 64 * 1024;   // LDS budget with staged weights
constexpr size_t kZoomLdsMax = 156 * 1024;    // dynamic LDS; k_zoom's static arrays take the rest
constexpr int kZoomRowTiles = 8;               // H chunk: at most 8 row tiles (128 rows)

// row tiles of the screen's H chunk: at most kZoomRowTiles and what fits
// beside the rest, balanced over the chunks (0: the zoom does not fit)
static int zoom_row_tiles(int R, int zf, bool stage_d)
{
    const int ZT = (zf * R + 15) / 16;
    const ZoomLayout base = zoom_layout(R, zf, stage_d, 0);
    if (base.total >= kZoomLdsMax) return 0;
    const size_t per = 2 * 16 * (size_t)((R + 31) / 32 * 32 + 8);
    const int fit = (int)std::min<size_t>(kZoomRowTiles, (kZoomLdsMax - base.total) / per);
    if (fit < 1) return 0;
    const int chunks = (ZT + fit - 1) / fit;
    return (ZT + chunks - 1) / chunks;
}

bool zoom_fits(int R, int zf) { return R <= 160 && zoom_row_tiles(R, zf, false) > 0; }

int zoom_fragment_count(int R, int zf) { return (zf * R + 15) / 16 * ((R + 31) / 32) * kWave * 8; }

hipError_t launch_zoom(hipStream_t s, const float *gamma_refined, int pairs, int R, int zf,
                       const double *dirichlet, const uint16_t *dfrag, const uint16_t *pfrag,
                       const ZoomBounds &bounds, const int2 *window, const float *coarse_peak,
                       int margin, int rf, float *offsets, float *refined_peak, int2 *zoom_loc,
                       const ZoomScratch &scratch, unsigned *stats)
{
    const double ecoef = bounds.ecoef, eabs = bounds.eabs;
    if (!zoom_fits(R, zf) || !scratch.thr || !scratch.win || !scratch.flag || !scratch.ncand ||
        !scratch.cand)
        return hipErrorInvalidValue;
    if (pairs == 0) return hipSuccess;
    if (stats) {
        hipLaunchKernelGGL(k_guard_clear, dim3(1), dim3(64), 0, s, stats);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const ZoomHandoff ho{scratch.thr, reinterpret_cast<int4 *>(scratch.win), scratch.flag,
                         scratch.ncand, scratch.cand};
    const auto *df = reinterpret_cast<const uint4 *>(dfrag);
    // the batched screen keeps a wave's fragments (NT tiles x KC chunks) in registers
    const int ZT = (zf * R + 15) / 16, KC = (R + 31) / 32;
    const int NT = (ZT + 3) / 4;
    const int nt = NT <= 2 ? 2 : NT <= 3 ? 3 : NT <= 5 ? 5 : NT <= 9 ? 9 : 0;
    // one wave per pair where both fragment tables fit in LDS, else the block screen
    const bool screen1 = ZT <= kScreenMaxTiles && KC <= 2 && zoom_screen1_lds(R, zf) <= kZoomLdsMax;
    const bool screen = screen1 || (nt > 0 && nt * KC <= 24 &&
                                    zoom_screen_lds(R, zf) <= kZoomLdsMax &&
                                    R * R <= kZoomScreenLoads * kZoomThreads);
    // 1. the exact window
    {
        const bool stage_d = zoom_layout(R, zf, true, 0).gt <= kZoomLdsStage;
        const size_t lds = zoom_layout(R, zf, stage_d, 0).gt;
        if (stage_d)
            hipLaunchKernelGGL(k_zoom_window<true>, dim3(pairs), dim3(kZoomThreads), lds, s,
                               gamma_refined, R, zf, dirichlet, ecoef, eabs, screen ? 0 : 1,
                               window, coarse_peak, margin, rf, offsets, refined_peak, zoom_loc, ho);
        else
            hipLaunchKernelGGL(k_zoom_window<false>, dim3(pairs), dim3(kZoomThreads), lds, s,
                               gamma_refined, R, zf, dirichlet, ecoef, eabs, screen ? 0 : 1,
                               window, coarse_peak, margin, rf, offsets, refined_peak, zoom_loc, ho);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // 2. the batched screen
    if (screen1) {
        const int blocks = std::min((pairs + kScreenWaves - 1) / kScreenWaves, 2 * 256);
        const size_t lds = zoom_screen1_lds(R, zf);
        const auto *pf = reinterpret_cast<const uint4 *>(pfrag);
        if (KC == 1)
            hipLaunchKernelGGL((k_zoom_screen1<1>), dim3(blocks), dim3(kScreenWaves * kWave), lds,
                               s, gamma_refined, pairs, R, zf, df, pf, bounds.lambda,
                               bounds.hcoef, bounds.habs, ho);
        else
            hipLaunchKernelGGL((k_zoom_screen1<2>), dim3(blocks), dim3(kScreenWaves * kWave), lds,
                               s, gamma_refined, pairs, R, zf, df, pf, bounds.lambda,
                               bounds.hcoef, bounds.habs, ho);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    } else if (screen) {
        const dim3 grid((pairs + kZoomScreenPairs - 1) / kZoomScreenPairs);
        const size_t lds = zoom_screen_lds(R, zf);
#define AMPCOR_SCREEN(KCV, NTV)                                                                  \
    hipLaunchKernelGGL((k_zoom_screen<KCV, NTV>), grid, dim3(kZoomThreads), lds, s, gamma_refined, \
                       pairs, R, zf, df, ho)
#define AMPCOR_SCREEN_NT(KCV)                               \
    switch (nt) {                                           \
    case 2: AMPCOR_SCREEN(KCV, 2); break;                   \
    case 3: AMPCOR_SCREEN(KCV, 3); break;                   \
    case 5: AMPCOR_SCREEN(KCV, 5); break;                   \
    default: AMPCOR_SCREEN(KCV, 9); break;                  \
    }
        switch (KC) {
        case 1: AMPCOR_SCREEN_NT(1) break;
        case 2: AMPCOR_SCREEN_NT(2) break;
        case 3: AMPCOR_SCREEN(3, 5); break;   // nt * KC <= 24: nt <= 5 here
        case 4: AMPCOR_SCREEN(4, 5); break;
        default: AMPCOR_SCREEN(5, 3); break;
        }
#undef AMPCOR_SCREEN_NT
#undef AMPCOR_SCREEN
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // 3. the listed cells and the edge maxima of the other pairs (shared rows:
    // up to kZoomSharedRows, as many as fit beside the weights and the surface)
    {
        const bool stage_d = zoom_cands_lds(R, zf, true, 0) <= kZoomLdsStage;
        const size_t base = zoom_cands_lds(R, zf, stage_d, 0);
        const int nr = base >= kZoomLdsStage
                           ? 0
                           : (int)std::min<size_t>(kZoomSharedRows,
                                                   (kZoomLdsStage - base) / (sizeof(double) * R));
        const size_t lds = zoom_cands_lds(R, zf, stage_d, nr);
        if (lds > kZoomLdsMax) return hipErrorInvalidValue;
        const int blocks = std::min(pairs, kZoomCandBlocks);
        if (stage_d)
            hipLaunchKernelGGL(k_zoom_cands<true>, dim3(blocks), dim3(kZoomThreads), lds, s,
                               gamma_refined, pairs, R, zf, dirichlet, nr, window, coarse_peak,
                               margin, rf, offsets, refined_peak, zoom_loc, stats, ho);
        else
            hipLaunchKernelGGL(k_zoom_cands<false>, dim3(blocks), dim3(kZoomThreads), lds, s,
                               gamma_refined, pairs, R, zf, dirichlet, nr, window, coarse_peak,
                               margin, rf, offsets, refined_peak, zoom_loc, stats, ho);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // 4. the full search of the flagged pairs
    const bool stage_d = zoom_layout(R, zf, true, zoom_row_tiles(R, zf, true)).total <=
                         kZoomLdsStage;
    const int rt = zoom_row_tiles(R, zf, stage_d);
    const size_t lds = zoom_layout(R, zf, stage_d, rt).total;
#define AMPCOR_ZOOM(ST, KCV)                                                                     \
    hipLaunchKernelGGL((k_zoom_full<ST, KCV>), dim3(pairs), dim3(kZoomThreads), lds, s,          \
                       gamma_refined, R, zf, dirichlet, df, ecoef, eabs, rt, window, coarse_peak, \
                       margin, rf, offsets, refined_peak, zoom_loc, stats, scratch.flag)
#define AMPCOR_ZOOM_KC(ST)                                  \
    switch (KC) {                                           \
    case 1: AMPCOR_ZOOM(ST, 1); break;                      \
    case 2: AMPCOR_ZOOM(ST, 2); break;                      \
    case 3: AMPCOR_ZOOM(ST, 3); break;                      \
    case 4: AMPCOR_ZOOM(ST, 4); break;                      \
    case 5: AMPCOR_ZOOM(ST, 5); break;                      \
    default: return hipErrorInvalidValue;                   \
    }
    if (stage_d) {
        AMPCOR_ZOOM_KC(true)
    } else {
        AMPCOR_ZOOM_KC(false)
    }
#undef AMPCOR_ZOOM_KC
#undef AMPCOR_ZOOM
    return hipGetLastError();
}

}  // namespace ampcor_hip
