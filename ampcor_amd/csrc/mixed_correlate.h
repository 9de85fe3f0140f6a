// mixed_correlate.h -- the contraction of the mixed-precision form on the
// matrix cores (v_mfma_f32_16x16x32 bf16/f16, FP32 accumulate).
//
// For chip rows i, i + 1 and a 16 x 16 tile of placements (u, v) one MFMA
// step with K = 32 = (two chip rows) x (16 target columns) computes
//   D[m][n] += sum_{d, x} T[16 ut + m + i + d][16 (V + kappa) + x]
//                         r[i + d][16 kappa + x - n]
// for every K-step kappa < NK = C / 16 + 1 (the 16-column blocks that a
// placement row's window crosses) and v-tile V < VT, so the target fragment
// of (V, kappa) is fragment s = V + kappa of NF = VT + NK - 1 sixteen-column
// fragments.  Lane l = 16 lk + m holds A[m][k = 8 lk + e] and
// B[k = 8 lk + e][n = m]; k = 16 d + 8 h + e with d = lk >> 1, h = lk & 1.
//
// Target fragments never leave registers: going from rows (i, i + 1) to
// (i + 2, i + 3) shifts every fragment up two rows, which is DPP row_shl:2
// inside each group of 16 lanes plus two new rows for lanes 14 and 15 (an
// exec-masked LDS read).  The first kReload fragments are read again from
// LDS instead, trading LDS bandwidth for the VALU time of the shifts.
// Toeplitz (chip) fragments are eight consecutive chip values starting at
// 16 kappa + 8 h - n (+16 padding); LDS holds each chip row twice, the
// second copy shifted by one element, so that this start is even in copy
// n & 1 and the fragment is four aligned dwords (two ds_read2_b32).
//
// The split forms add hi x lo and lo x hi products: the lo chip fragments
// are read like the hi ones, the lo target fragments straight from LDS.
//
// Included by the translation units that instantiate the launcher for a
// range of NK (the instantiations compile in parallel).
#pragma once

#include "kernels.h"

#include <algorithm>

namespace ampcor_hip {

namespace mixed_detail {

constexpr int kWave = 64;
constexpr int kMaxVT = 5;
// long chips (many K-steps) hold more fragments: at most four waves per
// workgroup, so that a wave may use the whole register file
constexpr int kWideNK = 12;
constexpr int max_waves(int nk) { return nk >= kWideNK ? 4 : 8; }
constexpr int kStageBatch = 8;  // global loads in flight per thread while staging

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

template <bool FP16>
__device__ inline floatx4 mfma32(uint4 a, uint4 b, floatx4 c)
{
    if constexpr (FP16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(halfx8, a),
                                                      __builtin_bit_cast(halfx8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// lanes 0..13 of each group of 16 take lane + 2's value, lanes 14, 15 keep `incoming`
__device__ inline uint32_t shift_up2(uint32_t cur, uint32_t incoming)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)incoming, (int)cur, 0x102, 0xf, 0xf, false);
}

__device__ inline uint4 shift_up2(uint4 cur, uint4 incoming)
{
    return make_uint4(shift_up2(cur.x, incoming.x), shift_up2(cur.y, incoming.y),
                      shift_up2(cur.z, incoming.z), shift_up2(cur.w, incoming.w));
}

template <int NK, int VT>
struct Geometry {
    static constexpr int CS = 16 * NK + 16;   // chip row: 16 zeros, C values, zeros
    static constexpr int NF = VT + NK - 1;    // target fragments per row
    static constexpr int ncols = 16 * NF;     // staged target columns
    static constexpr int TS = ncols + 8;      // staged row stride (elements), 16-byte aligned
    static constexpr int kReload = NF > 4 ? 4 : 0;
};

template <int NK, int VT, bool FP16, bool SPLIT>
__global__ __launch_bounds__(64 * max_waves(NK)) void k_correlate_mfma(
    const uint16_t *__restrict__ chip_q, const uint16_t *__restrict__ tgt_q, int C, int W, int P,
    int wg_ut, int tiles_u, int tiles_v, int nsplit, float *__restrict__ partial,
    const uint16_t *__restrict__ chip_lo, const uint16_t *__restrict__ tgt_lo)
{
    using G = Geometry<NK, VT>;
    constexpr int CS = G::CS, NF = G::NF, ncols = G::ncols, TS = G::TS, R = G::kReload;
    constexpr int planes = SPLIT ? 2 : 1;
    extern __shared__ uint4 lds_u4[];
    uint16_t *lds = reinterpret_cast<uint16_t *>(lds_u4);
    int bidx = xcd_block(blockIdx.x, gridDim.x);
    const int part = bidx % nsplit;
    bidx /= nsplit;
    const int tv = bidx % tiles_v;
    bidx /= tiles_v;
    const int tu = bidx % tiles_u;
    const int pair = bidx / tiles_u;

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int ut0 = tu * wg_ut, vt0 = tv * VT;
    const int ilen = C / nsplit, i0 = part * ilen;
    const int nrows = 16 * wg_ut + ilen - 1;
    uint16_t *tile = lds;                                  // [plane][nrows][TS]
    uint16_t *chip = lds + (size_t)planes * nrows * TS;    // [plane][ilen][2][CS]

    // stage the target block (zero outside the window) and the chip rows;
    // kStageBatch loads in flight per thread
    const int r0 = 16 * ut0 + i0, c0 = 16 * vt0;
    const bool wide = (W & 7) == 0;
#pragma unroll
    for (int pl = 0; pl < planes; ++pl) {
        const uint16_t *src = (pl ? tgt_lo : tgt_q) + (size_t)pair * W * W;
        uint16_t *dstt = tile + (size_t)pl * nrows * TS;
        const int total = nrows * (ncols / 8);
        for (int k0 = threadIdx.x; k0 < total; k0 += kStageBatch * blockDim.x) {
            uint4 v[kStageBatch];
#pragma unroll
            for (int j = 0; j < kStageBatch; ++j) {
                const int k = k0 + j * blockDim.x;
                const int rr = k / (ncols / 8), cc = 8 * (k - rr * (ncols / 8));
                const int gr = r0 + rr, gc = c0 + cc;
                v[j] = make_uint4(0u, 0u, 0u, 0u);
                if (k < total && gr < W && gc < W) {
                    const uint16_t *row = src + (size_t)gr * W;
                    if (wide && gc + 8 <= W) {
                        v[j] = *reinterpret_cast<const uint4 *>(row + gc);
                    } else {
                        uint32_t e[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) e[q] = (gc + q < W) ? row[gc + q] : 0u;
                        v[j] = make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16),
                                          e[4] | (e[5] << 16), e[6] | (e[7] << 16));
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < kStageBatch; ++j) {
                const int k = k0 + j * blockDim.x;
                if (k < total) {
                    const int rr = k / (ncols / 8), cc = 8 * (k - rr * (ncols / 8));
                    *reinterpret_cast<uint4 *>(dstt + (size_t)rr * TS + cc) = v[j];
                }
            }
        }
        const uint16_t *csrc = (pl ? chip_lo : chip_q) + ((size_t)pair * C + i0) * CS;
        uint16_t *dstc = chip + (size_t)pl * ilen * 2 * CS;
        const int ctotal = ilen * (CS / 4);
        for (int k0 = threadIdx.x; k0 < ctotal; k0 += kStageBatch * blockDim.x) {
            uint2 v[kStageBatch];
            uint32_t prev[kStageBatch];
#pragma unroll
            for (int j = 0; j < kStageBatch; ++j) {
                const int k = k0 + j * blockDim.x;
                v[j] = make_uint2(0u, 0u);
                prev[j] = 0u;
                if (k < ctotal) {
                    const int i = k / (CS / 4), c = 4 * (k - i * (CS / 4));
                    const uint16_t *rowp = csrc + (size_t)i * CS;
                    v[j] = *reinterpret_cast<const uint2 *>(rowp + c);
                    prev[j] = c ? (uint32_t)rowp[c - 1] : 0u;
                }
            }
#pragma unroll
            for (int j = 0; j < kStageBatch; ++j) {
                const int k = k0 + j * blockDim.x;
                if (k < ctotal) {
                    const int i = k / (CS / 4), c = 4 * (k - i * (CS / 4));
                    uint16_t *d0 = dstc + (size_t)i * 2 * CS + c;
                    *reinterpret_cast<uint2 *>(d0) = v[j];
                    *reinterpret_cast<uint2 *>(d0 + CS) =
                        make_uint2(prev[j] | (v[j].x << 16), (v[j].x >> 16) | (v[j].y << 16));
                }
            }
        }
    }
    __syncthreads();

    const int ut = ut0 + wave;
    if (wave >= wg_ut || 16 * ut >= P) return;

    const int d = lk >> 1, h = lk & 1;   // chip row of the pair, half of the 16 columns
    // chip fragment of K-step kappa: dwords 8 kappa + 4 h + 8 - (lm >> 1) .. +3
    // of copy (lm & 1) of chip row i + d
    const uint32_t *cw = reinterpret_cast<const uint32_t *>(chip) + d * CS + (lm & 1) * (CS / 2) +
                         4 * h + 8 - (lm >> 1);
    const uint16_t *tw = tile + (size_t)(16 * wave + d) * TS + 8 * h;
    const int last = nrows - 1 - 16 * wave - d;   // last row this lane may read, relative to tw

    uint4 A[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f)
        A[f] = *reinterpret_cast<const uint4 *>(tw + (size_t)lm * TS + 16 * f);

    // two accumulators per tile (even and odd K-steps): twice the independent
    // MFMA chains; summed at the end
    floatx4 acc[VT], acc2[VT];
#pragma unroll
    for (int v = 0; v < VT; ++v) acc[v] = acc2[v] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    auto step = [&](const int ii) {
        const uint32_t *cr = cw + (size_t)ii * CS;
        uint4 B[NK], BL[SPLIT ? NK : 1], AL[SPLIT ? NF : 1];
#pragma unroll
        for (int kap = 0; kap < NK; ++kap)
            B[kap] = make_uint4(cr[8 * kap], cr[8 * kap + 1], cr[8 * kap + 2], cr[8 * kap + 3]);
        if constexpr (SPLIT) {
            const uint32_t *crl = cr + (size_t)ilen * CS;
#pragma unroll
            for (int kap = 0; kap < NK; ++kap)
                BL[kap] = make_uint4(crl[8 * kap], crl[8 * kap + 1], crl[8 * kap + 2],
                                     crl[8 * kap + 3]);
            const uint16_t *twl = tw + (size_t)nrows * TS + (size_t)(lm + ii) * TS;
#pragma unroll
            for (int f = 0; f < NF; ++f) AL[f] = *reinterpret_cast<const uint4 *>(twl + 16 * f);
        }
        // rows entering at the next step (the last step reads rows it does not
        // use, clamped into the block)
        const int nr = min(lm + 2 + ii, last);
        uint4 Nx[NF], Rl[R > 0 ? R : 1];
        if (lm >= 14) {
#pragma unroll
            for (int f = R; f < NF; ++f)
                Nx[f] = *reinterpret_cast<const uint4 *>(tw + (size_t)nr * TS + 16 * f);
        }
#pragma unroll
        for (int f = 0; f < R; ++f) Rl[f] = *reinterpret_cast<const uint4 *>(tw + (size_t)nr * TS + 16 * f);
#pragma unroll
        for (int kap = 0; kap < NK; ++kap)
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                floatx4 &c = (kap & 1) ? acc2[v] : acc[v];
                c = mfma32<FP16>(A[v + kap], B[kap], c);
                if constexpr (SPLIT) {
                    c = mfma32<FP16>(A[v + kap], BL[kap], c);
                    c = mfma32<FP16>(AL[v + kap], B[kap], c);
                }
            }
#pragma unroll
        for (int f = 0; f < R; ++f) A[f] = Rl[f];
#pragma unroll
        for (int f = R; f < NF; ++f) A[f] = shift_up2(A[f], Nx[f]);
    };
    // two steps (four chip rows) per trip: the shifted fragments land in the
    // registers of the incoming rows and return to the first set
    int ii = 0;
    for (; ii + 3 < ilen; ii += 4) {
        step(ii);
        step(ii + 2);
    }
    if (ii < ilen) step(ii);

    float *dst = partial + ((size_t)pair * nsplit + part) * P * P;
#pragma unroll
    for (int v = 0; v < VT; ++v) {
        const int vv = 16 * (vt0 + v) + lm;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int uu = 16 * ut + 4 * lk + e;
            if (uu < P && vv < P) dst[(size_t)uu * P + vv] = acc[v][e] + acc2[v][e];
        }
    }
}

// ---- the streaming form -------------------------------------------------------
//
// One workgroup per (pair, block of u- and v-tiles) over all C chip rows, which
// pass through LDS kStreamRows at a time: while the waves multiply one chunk,
// each thread's global loads for the next chunk are in flight, and their data
// go to LDS slots the current chunk does not read, so one barrier per chunk
// suffices.  Target rows live in a ring of RS = 16 wg_ut + 2 kStreamRows rows
// (the rows of the current chunk and the next chunk's new rows); the chip rows
// of a chunk take one of two slots.  The fragments, the DPP row shifts and the
// MFMA steps are those of k_correlate_mfma; the shift that crosses into the
// next chunk waits for the barrier.  All chip rows in one workgroup: one
// partial surface per pair.
constexpr int kStreamRows = 8;   // chip rows per chunk
constexpr int kStreamItems = 4;  // 16-byte loads per thread per chunk, at most

template <int NK, int VT>
struct StreamGeometry {
    using G = Geometry<NK, VT>;
    static constexpr int CPR = G::ncols / 8;   // 16-byte chunks per staged target row
    static constexpr int CCH = G::CS / 8;      // 16-byte chunks per chip row
    // loads of one chunk: kStreamRows new target rows and kStreamRows chip rows
    static constexpr int items = kStreamRows * (CPR + CCH);
};

template <int NK, int VT, bool FP16>
__global__ __launch_bounds__(64 * max_waves(NK)) void k_correlate_stream(
    const uint16_t *__restrict__ chip_q, const uint16_t *__restrict__ tgt_q, int C, int W, int P,
    int wg_ut, int tiles_u, int tiles_v, float *__restrict__ partial)
{
    using G = Geometry<NK, VT>;
    using SG = StreamGeometry<NK, VT>;
    constexpr int CS = G::CS, NF = G::NF, TS = G::TS, R = G::kReload;
    constexpr int CPR = SG::CPR, CCH = SG::CCH, CH = kStreamRows;
    extern __shared__ uint4 lds_u4[];
    uint16_t *lds = reinterpret_cast<uint16_t *>(lds_u4);
    int bidx = xcd_block(blockIdx.x, gridDim.x);
    const int tv = bidx % tiles_v;
    bidx /= tiles_v;
    const int tu = bidx % tiles_u;
    const int pair = bidx / tiles_u;

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
    const int lm = lane & 15, lk = lane >> 4;
    const int ut0 = tu * wg_ut, vt0 = tv * VT;
    const int RS = 16 * wg_ut + 2 * CH;            // target ring rows
    uint16_t *ring = lds;                          // [RS][TS]
    uint16_t *chip = lds + (size_t)RS * TS;        // [2 slots][CH rows][2 copies][CS]
    const int r0 = 16 * ut0, c0 = 16 * vt0;        // block origin in the window
    const uint16_t *src = tgt_q + (size_t)pair * W * W;
    const uint16_t *csrc = chip_q + (size_t)pair * C * CS;
    const int nchunks = C / CH;
    const int first_rows = 16 * wg_ut + CH - 1;    // target rows of chunk 0

    // item k of the loads that bring chunk c (c >= 1): its new target rows,
    // then its chip rows; `prev` is the dword before a chip chunk (copy 1's
    // first element), zero at the start of a row
    auto load_item = [&](const int c, const int k, uint4 &v, uint32_t &prev) {
        v = make_uint4(0u, 0u, 0u, 0u);
        prev = 0u;
        if (k < CH * CPR) {
            const int rr = k / CPR, cc = 8 * (k - rr * CPR);
            const int gr = r0 + first_rows + CH * (c - 1) + rr, gc = c0 + cc;
            const bool in = gr < W && gc < W;
            const uint4 x = *reinterpret_cast<const uint4 *>(src + (in ? (size_t)gr * W + gc : 0));
            if (in) v = x;
        } else if (k < SG::items) {
            const int kc = k - CH * CPR, i = kc / CCH, q = kc - i * CCH;
            const uint16_t *row = csrc + (size_t)(CH * c + i) * CS;
            v = *reinterpret_cast<const uint4 *>(row + 8 * q);
            if (q) prev = *reinterpret_cast<const uint32_t *>(row + 8 * q - 2);
        }
    };
    auto store_chip = [&](const int slot, const int i, const int q, const uint4 v,
                          const uint32_t prev) {
        uint16_t *d0 = chip + ((size_t)(slot * CH + i) * 2) * CS + 8 * q;
        *reinterpret_cast<uint4 *>(d0) = v;
        *reinterpret_cast<uint4 *>(d0 + CS) =
            make_uint4((prev >> 16) | (v.x << 16), (v.x >> 16) | (v.y << 16),
                       (v.y >> 16) | (v.z << 16), (v.z >> 16) | (v.w << 16));
    };
    auto store_item = [&](const int c, const int k, const uint4 v, const uint32_t prev) {
        if (k < CH * CPR) {
            const int rr = k / CPR, cc = 8 * (k - rr * CPR);
            const int lr = (first_rows + CH * (c - 1) + rr) % RS;
            *reinterpret_cast<uint4 *>(ring + (size_t)lr * TS + cc) = v;
        } else if (k < SG::items) {
            const int kc = k - CH * CPR, i = kc / CCH;
            store_chip(c & 1, i, kc - i * CCH, v, prev);
        }
    };

    // chunk 0: its target rows and chip rows
    for (int k = threadIdx.x; k < first_rows * CPR; k += blockDim.x) {
        const int rr = k / CPR, cc = 8 * (k - rr * CPR);
        const int gr = r0 + rr, gc = c0 + cc;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gr < W && gc < W) v = *reinterpret_cast<const uint4 *>(src + (size_t)gr * W + gc);
        *reinterpret_cast<uint4 *>(ring + (size_t)rr * TS + cc) = v;
    }
    for (int kc = threadIdx.x; kc < CH * CCH; kc += blockDim.x) {
        const int i = kc / CCH, q = kc - i * CCH;
        const uint16_t *row = csrc + (size_t)i * CS;
        store_chip(0, i, q, *reinterpret_cast<const uint4 *>(row + 8 * q),
                   q ? *reinterpret_cast<const uint32_t *>(row + 8 * q - 2) : 0u);
    }
    __syncthreads();

    const int ut = ut0 + wave;
    // waves past the surface still load and meet every barrier
    const bool active = wave < wg_ut && 16 * ut < P;
    const int d = lk >> 1, h = lk & 1;
    // this lane's target row (ring index) at the current step
    int row = 16 * wave + d + lm;
    const uint16_t *twc = ring + 8 * h;            // column origin of this lane's fragments
    uint4 A[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f)
        A[f] = *reinterpret_cast<const uint4 *>(twc + (size_t)row * TS + 16 * f);

    floatx4 acc[VT], acc2[VT];
#pragma unroll
    for (int v = 0; v < VT; ++v) acc[v] = acc2[v] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    // the shift from one step to the next: rows two further down
    auto advance = [&]() {
        int nr = row + 2;
        nr -= nr >= RS ? RS : 0;
        uint4 Nx[NF], Rl[R > 0 ? R : 1];
        if (lm >= 14) {
#pragma unroll
            for (int f = R; f < NF; ++f)
                Nx[f] = *reinterpret_cast<const uint4 *>(twc + (size_t)nr * TS + 16 * f);
        }
#pragma unroll
        for (int f = 0; f < R; ++f)
            Rl[f] = *reinterpret_cast<const uint4 *>(twc + (size_t)nr * TS + 16 * f);
#pragma unroll
        for (int f = 0; f < R; ++f) A[f] = Rl[f];
#pragma unroll
        for (int f = R; f < NF; ++f) A[f] = shift_up2(A[f], Nx[f]);
        row = nr;
    };
    // the MFMAs of one pair of chip rows (cr: this lane's row d of them)
    auto multiply = [&](const uint32_t *cr) {
        uint4 B[NK];
#pragma unroll
        for (int kap = 0; kap < NK; ++kap)
            B[kap] = make_uint4(cr[8 * kap], cr[8 * kap + 1], cr[8 * kap + 2], cr[8 * kap + 3]);
#pragma unroll
        for (int kap = 0; kap < NK; ++kap)
#pragma unroll
            for (int v = 0; v < VT; ++v) {
                floatx4 &c = (kap & 1) ? acc2[v] : acc[v];
                c = mfma32<FP16>(A[v + kap], B[kap], c);
            }
    };
    // chip fragment of K-step kappa: dwords 8 kappa + 4 h + 8 - (lm >> 1) .. +3
    // of copy (lm & 1) of chip row d of the step
    const uint32_t *cw = reinterpret_cast<const uint32_t *>(chip) + d * CS + (lm & 1) * (CS / 2) +
                         4 * h + 8 - (lm >> 1);

    for (int c = 0; c < nchunks; ++c) {
        const bool more = c + 1 < nchunks;
        uint4 nv[kStreamItems];
        uint32_t np[kStreamItems];
        if (more) {
#pragma unroll
            for (int j = 0; j < kStreamItems; ++j)
                load_item(c + 1, j * (int)blockDim.x + (int)threadIdx.x, nv[j], np[j]);
        }
        if (active) {
            const uint32_t *slot = cw + (size_t)(c & 1) * CH * CS;
#pragma unroll
            for (int sr = 0; sr < CH; sr += 2) {
                multiply(slot + (size_t)sr * CS);
                if (sr + 2 < CH) advance();
            }
        }
        if (more) {
#pragma unroll
            for (int j = 0; j < kStreamItems; ++j)
                store_item(c + 1, j * (int)blockDim.x + (int)threadIdx.x, nv[j], np[j]);
        }
        __syncthreads();
        if (active && more) advance();
    }
    if (!active) return;

    float *dst = partial + (size_t)pair * P * P;
#pragma unroll
    for (int v = 0; v < VT; ++v) {
        const int vv = 16 * (vt0 + v) + lm;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int uu = 16 * ut + 4 * lk + e;
            if (uu < P && vv < P) dst[(size_t)uu * P + vv] = acc[v][e] + acc2[v][e];
        }
    }
}

}  // namespace mixed_detail

// LDS bytes of one workgroup
inline size_t mixed_lds_bytes(int C, int nk, int vt, int wg_ut, int nsplit, bool split)
{
    const int ilen = C / nsplit, cs = 16 * nk + 16;
    const size_t rows = 16 * wg_ut + ilen - 1, cols = 16 * (vt + nk - 1) + 8;
    return (rows * cols + (size_t)ilen * 2 * cs) * sizeof(uint16_t) * (split ? 2 : 1);
}

// the streaming form: LDS bytes of one workgroup (target ring and two chip
// slots), and the 16-byte loads that bring one chunk
inline size_t mixed_stream_lds_bytes(int nk, int vt, int wg_ut)
{
    using namespace mixed_detail;
    const size_t ring = 16 * wg_ut + 2 * kStreamRows, cols = 16 * (vt + nk - 1) + 8;
    return (ring * cols + (size_t)2 * kStreamRows * 2 * (16 * nk + 16)) * sizeof(uint16_t);
}
inline int mixed_stream_items(int nk, int vt)
{
    return mixed_detail::kStreamRows * (2 * (vt + nk - 1) + 2 * nk + 2);
}

template <int NK>
hipError_t launch_correlate_mixed_nk(hipStream_t s, const MixedShape &sh, const uint16_t *chip_q,
                                     const uint16_t *tgt_q, int pairs, int C, int W, bool fp16,
                                     float *partial, const uint16_t *chip_lo,
                                     const uint16_t *tgt_lo)
#ifdef AMPCOR_MIXED_INSTANTIATE
{
    using namespace mixed_detail;
    const int P = W - C + 1;
    const size_t per = (size_t)sh.tiles_u * sh.tiles_v * sh.nsplit;   // workgroups per pair
    const dim3 block(64 * sh.wg_ut);
    // a launch's work-items fit in 32 bits: pairs go in chunks
    const int chunk = std::min(pairs, launch_pair_cap(per * block.x));
#define AMPCOR_MIXED_ONE(VTV, F16, SPL)                                                            \
    hipLaunchKernelGGL((k_correlate_mfma<NK, VTV, F16, SPL>), grid, block, sh.lds_bytes, s, cq,   \
                       tq, C, W, P, sh.wg_ut, sh.tiles_u, sh.tiles_v, sh.nsplit, pt, cl, tl)
#define AMPCOR_MIXED_VT(VTV)                                                                       \
    do {                                                                                           \
        if (sh.split) {                                                                            \
            if (fp16) AMPCOR_MIXED_ONE(VTV, true, true);                                           \
            else AMPCOR_MIXED_ONE(VTV, false, true);                                               \
        } else {                                                                                   \
            if (fp16) AMPCOR_MIXED_ONE(VTV, true, false);                                          \
            else AMPCOR_MIXED_ONE(VTV, false, false);                                              \
        }                                                                                          \
    } while (0)
    if (sh.stream) {
        for (int first = 0; first < pairs; first += chunk) {
            const int n = std::min(chunk, pairs - first);
            const dim3 grid((unsigned)(n * per));
            const uint16_t *cq = chip_q + (size_t)first * C * sh.cs;
            const uint16_t *tq = tgt_q + (size_t)first * W * W;
            float *pt = partial + (size_t)first * P * P;
#define AMPCOR_STREAM(VTV)                                                                         \
    case VTV:                                                                                      \
        if (fp16)                                                                                  \
            hipLaunchKernelGGL((k_correlate_stream<NK, VTV, true>), grid, block, sh.lds_bytes, s,   \
                               cq, tq, C, W, P, sh.wg_ut, sh.tiles_u, sh.tiles_v, pt);             \
        else                                                                                       \
            hipLaunchKernelGGL((k_correlate_stream<NK, VTV, false>), grid, block, sh.lds_bytes, s,  \
                               cq, tq, C, W, P, sh.wg_ut, sh.tiles_u, sh.tiles_v, pt);             \
        break
            switch (sh.wg_vt) {
                AMPCOR_STREAM(1);
                AMPCOR_STREAM(2);
                AMPCOR_STREAM(3);
                AMPCOR_STREAM(4);
                AMPCOR_STREAM(5);
            default: return hipErrorInvalidValue;
            }
#undef AMPCOR_STREAM
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    for (int first = 0; first < pairs; first += chunk) {
        const int n = std::min(chunk, pairs - first);
        const dim3 grid((unsigned)(n * per));
        const size_t oc = (size_t)first * C * sh.cs, ot = (size_t)first * W * W;
        const uint16_t *cq = chip_q + oc, *tq = tgt_q + ot;
        const uint16_t *cl = chip_lo ? chip_lo + oc : nullptr;
        const uint16_t *tl = tgt_lo ? tgt_lo + ot : nullptr;
        float *pt = partial + (size_t)first * sh.nsplit * P * P;
        switch (sh.wg_vt) {
        case 1: AMPCOR_MIXED_VT(1); break;
        case 2: AMPCOR_MIXED_VT(2); break;
        case 3: AMPCOR_MIXED_VT(3); break;
        case 4: AMPCOR_MIXED_VT(4); break;
        case 5: AMPCOR_MIXED_VT(5); break;
        default: return hipErrorInvalidValue;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
#undef AMPCOR_MIXED_VT
#undef AMPCOR_MIXED_ONE
    return hipSuccess;
}
#else
    ;
#endif

}  // namespace ampcor_hip
