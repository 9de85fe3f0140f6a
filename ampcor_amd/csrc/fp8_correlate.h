// fp8_correlate.h -- the coarse contraction of the FP8 form on the matrix
// cores (v_mfma_f32_16x16x128_f8f6f4 with both operands OCP e4m3, FP32
// accumulate; twice the rate of the 16x16x32 BF16 instruction).
//
// One MFMA step covers eight chip rows i .. i + 7 and a 16 x 16 tile of
// placements (u, v): with K = 128 = (eight chip rows) x (16 target columns),
//   D[m][n] += sum_{d < 8, x < 16} T[16 ut + m + i + d][16 (V + kappa) + x]
//                                  r[i + d][16 kappa + x - n]
// for every K-step kappa < NK = C / 16 + 1 and v-tile V < VT, exactly as the
// BF16/FP16 contraction with two chip rows per step.  Lane l = 16 g + m holds
// 32 bytes of each operand: bytes j < 16 belong to chip row 2 g, bytes j >= 16
// to chip row 2 g + 1, column x = j mod 16 (the instruction pairs the bytes of
// A and B position by position, rows m = l mod 16 of A with columns n = l mod
// 16 of B, and returns D[4 g + e][l mod 16] in register e).
//
// Target fragments are two 16-byte row segments read from LDS for every step
// (no register shifting: one K=128 instruction does four times the work of a
// K=32 one per fragment read).  Toeplitz (chip) fragments start at byte
// 16 kappa + 16 - n of a chip row stored with 16 leading zero bytes; LDS holds
// each chip row four times, copy c shifted right by c bytes, so that in copy
// n mod 4 the start is a dword boundary and a fragment is four dword reads.
//
// Included by the translation unit that instantiates the launcher.
#pragma once

#include "kernels.h"

namespace ampcor_hip {

namespace fp8_detail {

constexpr int kWave = 64;
constexpr int kMaxVT = 5;
constexpr int kMaxWaves = 8;
constexpr int kStageBatch = 8;   // global loads in flight per thread while staging

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx8 __attribute__((ext_vector_type(8)));

template <int NK, int VT>
struct Geometry8 {
    static constexpr int CS = 16 * NK + 16;   // chip row bytes: 16 zeros, C values, zeros
    static constexpr int CS4 = CS + 4;        // one LDS copy (shifted by up to 3 bytes)
    static constexpr int NF = VT + NK - 1;    // target fragments per row
    static constexpr int ncols = 16 * NF;     // staged target columns (bytes)
    static constexpr int TS = ncols + 16;     // staged row stride (bytes), 16-byte aligned
};

__device__ inline floatx4 mfma_fp8(const intx8 &a, const intx8 &b, floatx4 c)
{
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, 0, 0, 0);
}

template <int NK, int VT>
__global__ __launch_bounds__(64 * kMaxWaves) void k_correlate_fp8(
    const uint8_t *__restrict__ chip_q, const uint8_t *__restrict__ tgt_q, int C, int W, int P,
    int ilen, int wg_ut, int tiles_u, int tiles_v, int nsplit, float *__restrict__ partial)
{
    using G = Geometry8<NK, VT>;
    constexpr int CS = G::CS, CS4 = G::CS4, NF = G::NF, ncols = G::ncols, TS = G::TS;
    extern __shared__ uint4 lds_u4[];
    uint8_t *lds = reinterpret_cast<uint8_t *>(lds_u4);
    int bidx = xcd_block(blockIdx.x, gridDim.x);
    const int part = bidx % nsplit;
    bidx /= nsplit;
    const int tv = bidx % tiles_v;
    bidx /= tiles_v;
    const int tu = bidx % tiles_u;
    const int pair = bidx / tiles_u;

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int lm = lane & 15, g = lane >> 4;
    const int ut0 = tu * wg_ut, vt0 = tv * VT;
    const int i0 = part * ilen;                        // ilen: a multiple of 8
    const int nrows = 16 * wg_ut + ilen - 1;
    const int tile_bytes = (nrows * TS + 15) / 16 * 16;
    uint8_t *tile = lds;                               // [nrows][TS]
    uint32_t *chip = reinterpret_cast<uint32_t *>(lds + tile_bytes);   // [ilen][4][CS4 / 4]

    // stage the target block (zero outside the window), 16 bytes per load
    {
        const uint8_t *src = tgt_q + (size_t)pair * W * W;
        const int r0 = 16 * ut0 + i0, c0 = 16 * vt0;
        const bool wide = (W & 15) == 0;
        constexpr int per_row = ncols / 16;
        const int total = nrows * per_row;
        for (int k0 = threadIdx.x; k0 < total; k0 += kStageBatch * blockDim.x) {
            uint4 v[kStageBatch];
#pragma unroll
            for (int j = 0; j < kStageBatch; ++j) {
                const int k = k0 + j * blockDim.x;
                const int rr = k / per_row, cc = 16 * (k - rr * per_row);
                const int gr = r0 + rr, gc = c0 + cc;
                v[j] = make_uint4(0u, 0u, 0u, 0u);
                if (k < total && gr < W && gc < W) {
                    const uint8_t *row = src + (size_t)gr * W;
                    if (wide) {
                        v[j] = *reinterpret_cast<const uint4 *>(row + gc);
                    } else {
                        uint32_t e[4] = {0u, 0u, 0u, 0u};
                        for (int q = 0; q < 16; ++q)
                            if (gc + q < W) e[q >> 2] |= (uint32_t)row[gc + q] << (8 * (q & 3));
                        v[j] = make_uint4(e[0], e[1], e[2], e[3]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < kStageBatch; ++j) {
                const int k = k0 + j * blockDim.x;
                if (k < total) {
                    const int rr = k / per_row, cc = 16 * (k - rr * per_row);
                    *reinterpret_cast<uint4 *>(tile + (size_t)rr * TS + cc) = v[j];
                }
            }
        }
    }
    // stage the chip rows i0 .. i0 + ilen - 1 (zero past C), four shifted copies:
    // dword w of copy c holds row bytes 4 w - c .. 4 w + 3 - c
    {
        const uint32_t *csrc = reinterpret_cast<const uint32_t *>(chip_q + (size_t)pair * C * CS);
        constexpr int wpr = CS4 / 4;   // dwords per copy
        const int total = ilen * wpr;
        for (int k = threadIdx.x; k < total; k += blockDim.x) {
            const int ii = k / wpr, w = k - ii * wpr;
            const int i = i0 + ii;
            uint32_t cur = 0u, prev = 0u;
            if (i < C) {
                const uint32_t *row = csrc + (size_t)i * (CS / 4);
                cur = w < CS / 4 ? row[w] : 0u;
                prev = w > 0 ? row[w - 1] : 0u;
            }
            uint32_t *dst = chip + (size_t)ii * 4 * wpr + w;
            dst[0] = cur;
            dst[wpr] = (cur << 8) | (prev >> 24);
            dst[2 * wpr] = (cur << 16) | (prev >> 16);
            dst[3 * wpr] = (cur << 24) | (prev >> 8);
        }
    }
    __syncthreads();

    const int ut = ut0 + wave;
    if (wave >= wg_ut || 16 * ut >= P) return;

    // target rows 16 wave + m + ii + 2 g (+1) of this lane
    const uint8_t *ta = tile + (size_t)(16 * wave + lm + 2 * g) * TS;
    // chip fragment of K-step kappa: dwords 4 kappa + 4 - (m >> 2) .. + 3 of
    // copy (m & 3) of chip rows ii + 2 g (+1)
    constexpr int wpr = CS4 / 4;
    const uint32_t *cb = chip + (size_t)(2 * g) * 4 * wpr + (lm & 3) * wpr + 4 - (lm >> 2);

    floatx4 acc[VT];
#pragma unroll
    for (int v = 0; v < VT; ++v) acc[v] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int ii = 0; ii < ilen; ii += 8) {
        intx8 B[NK];
        const uint32_t *c0 = cb + (size_t)ii * 4 * wpr, *c1 = c0 + 4 * wpr;
#pragma unroll
        for (int kap = 0; kap < NK; ++kap) {
            const uint32_t *p0 = c0 + 4 * kap, *p1 = c1 + 4 * kap;
            B[kap] = intx8{(int)p0[0], (int)p0[1], (int)p0[2], (int)p0[3],
                           (int)p1[0], (int)p1[1], (int)p1[2], (int)p1[3]};
        }
        const uint8_t *r0 = ta + (size_t)ii * TS, *r1 = r0 + TS;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const uint4 a0 = *reinterpret_cast<const uint4 *>(r0 + 16 * f);
            const uint4 a1 = *reinterpret_cast<const uint4 *>(r1 + 16 * f);
            const intx8 A = intx8{(int)a0.x, (int)a0.y, (int)a0.z, (int)a0.w,
                                  (int)a1.x, (int)a1.y, (int)a1.z, (int)a1.w};
#pragma unroll
            for (int kap = 0; kap < NK; ++kap) {
                const int v = f - kap;
                if (v >= 0 && v < VT) acc[v] = mfma_fp8(A, B[kap], acc[v]);
            }
        }
    }

    float *dst = partial + ((size_t)pair * nsplit + part) * P * P;
#pragma unroll
    for (int v = 0; v < VT; ++v) {
        const int vv = 16 * (vt0 + v) + lm;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int uu = 16 * ut + 4 * g + e;
            if (uu < P && vv < P) dst[(size_t)uu * P + vv] = acc[v][e];
        }
    }
}

}  // namespace fp8_detail

// LDS bytes of one workgroup
inline size_t fp8_lds_bytes(int nk, int vt, int wg_ut, int ilen)
{
    const size_t cs = 16 * nk + 16, rows = 16 * wg_ut + ilen - 1, ts = 16 * (vt + nk - 1) + 16;
    return (rows * ts + 15) / 16 * 16 + (size_t)ilen * 4 * (cs + 4);
}

}  // namespace ampcor_hip
