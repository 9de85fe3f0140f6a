// packed_correlate.h -- the contraction of the mixed-precision form for
// surfaces of P = 32 NB + 1 placements per axis (margins that are multiples of
// 16), with no padded placement rows.
//
// The MFMA step is that of k_correlate_mfma (v_mfma_f32_16x16x32, K = two chip
// rows x 16 target columns, the chip as a Toeplitz operand built from two
// shifted copies of each chip row in LDS), but the 16 M-rows of a tile are
// not 16 consecutive placement rows.  A wave owns the 32 rows of a u-block b
// as two interleaved tiles: M-row m of tile t is placement row
// u = 32 b + t + 2 m.  Going from chip rows (i, i + 1) to (i + 2, i + 3) moves
// every M-row's target row down by two, which is the row the next M-row of the
// same tile holds: each fragment dword takes one DPP row_shl:1, and only lane
// m = 15 takes a new row, which is the same row for every u-block's wave
// position: rows 32 (b + 1) + i + {0, 1, 2}.  Only those rows pass through LDS
// (a ring of 2 CH + 1 rows per u-block), and the chip rows through two slots of
// CH rows; both are refilled one chunk of CH chip rows ahead with one barrier
// per chunk.
//
// The last placement row, u = P - 1 = 32 NB, is a packed tile: its M-row m
// holds the segment of 16 placements that starts at v-tile m of the wave.  Its
// target row is the one that lane 15 of tile 0 loads anyway, so the same
// masked load brings both (lane 15: fragment f, lanes m < VT: fragment m + f).
// The last v-tile of a surface (v = P - 1 only) is an ordinary tile of which
// one column is kept.
//
// A workgroup is one pair, or one of its nsplit ranges of C / nsplit chip
// rows (each written to a partial surface of its own, which the normalizer
// sums in order): NB u-blocks x nvg groups of VT v-tiles, one wave each.
// Ranges give surfaces of one or two waves enough workgroups to spread evenly
// over the CUs.  Tiles per pair and chip-row pair: 2 NB nvg VT + nvg against
// (2 NB + 1)^2 for tiles of 16 consecutive rows (63 instead of 81 MFMAs per
// K-step column for P = 33 and one group of 3 v-tiles; 105 instead of 125 for
// P = 65).
//
// Included by the translation units that instantiate the launchers per NK.
#pragma once

#include "mixed_correlate.h"

#include <type_traits>

namespace ampcor_hip {

namespace packed_detail {

using mixed_detail::floatx4;
using mixed_detail::mfma32;
constexpr int kWave = 64;
constexpr int kMaxWaves = 12;   // at most three waves per SIMD (168 VGPRs)
// wave-wide refill loads per thread and step, at most: rows of 17 K-steps
// (one row per load) need twice as many
__host__ __device__ constexpr int refill_loads(int nk) { return nk > 11 ? 4 : 2; }
constexpr int kShift = 1;   // fragments shifted (and their lane-15 rows loaded) together
// dword 3 of a raw buffer resource on gfx9 targets (bounds-checked loads, zeros
// past num_records)
constexpr int kBufferWord3 = 0x00020000;

// lanes 0..14 of each group of 16 take lane + 1's value, lane 15 keeps `incoming`
__device__ inline uint32_t shift_up1(uint32_t cur, uint32_t incoming)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)incoming, (int)cur, 0x101, 0xf, 0xf, false);
}

__device__ inline uint4 shift_up1(uint4 cur, uint4 incoming)
{
    return make_uint4(shift_up1(cur.x, incoming.x), shift_up1(cur.y, incoming.y),
                      shift_up1(cur.z, incoming.z), shift_up1(cur.w, incoming.w));
}

// the LDS ring's rows hold the target rows of columns [16 g VT, ...) for every
// v-group g: 16 (nvg VT + NK - 1) columns, and 8 more so that rows fall on
// different banks
__host__ __device__ constexpr int packed_row_stride(int nk, int vt, int nvg)
{
    return 16 * (nvg * vt + nk - 1) + 8;
}

// WPE: the occupancy the register allocation is held to (waves per SIMD: 3
// leaves 168 VGPRs per lane, 2 leaves 256); a workgroup then has at most
// 4 WPE waves (one CU's four SIMDs)
template <int NK, int VT, int CH, int WPE, bool FP16>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 256 * WPE),
                          amdgpu_waves_per_eu(WPE, WPE))) void k_correlate_packed(
    const uint16_t *__restrict__ chip_q, const uint16_t *__restrict__ tgt_q, int C, int W, int P,
    int nvg, int nsplit, float *__restrict__ partial, int accumulate)
{
    constexpr int NF = VT + NK - 1;          // target fragments per tile
    constexpr int CS = 16 * NK + 16;         // chip row: 16 zeros, C values, zeros
    constexpr int RR = 2 * CH + 1;           // ring rows per u-block
    constexpr int CCH = CS / 8;              // 16-byte chunks per chip row
    extern __shared__ uint4 lds_u4[];
    uint16_t *lds = reinterpret_cast<uint16_t *>(lds_u4);

    // the ranges of a pair in consecutive workgroups (one XCD's L2 holds the window)
    const int wg = xcd_block(blockIdx.x, gridDim.x);
    const int pair = wg / nsplit, ks = wg - pair * nsplit;
    const int CR = C / nsplit, i0 = ks * CR;     // the range's chip rows [i0, i0 + CR)
    const int NB = (P - 1) >> 5;
    const int TS = packed_row_stride(NK, VT, nvg);
    const int TW = TS - 8;                   // staged columns per ring row
    const int CPR = TW / 8;                  // 16-byte chunks per ring row
    uint16_t *ring = lds;                                    // [NB][RR][TS]
    uint16_t *chip = lds + (size_t)NB * RR * TS;             // [2][CH][2][CS]
    // target row u + i of chip row i is row u + (i - i0) from src
    const uint16_t *src = tgt_q + (size_t)pair * W * W + (size_t)i0 * W;
    const uint16_t *csrc = chip_q + ((size_t)pair * C + i0) * CS;
    const int WR = W - i0;                       // window rows from src
    const int nchunks = CR / CH;

    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int lm = lane & 15, lk = lane >> 4;
    const int d = lk >> 1, h = lk & 1;
    const int b = wave / nvg, g = wave - b * nvg;        // wave-uniform
    const bool edge = b == NB - 1;           // this wave also holds the packed row u = P - 1
    const int col0 = 16 * g * VT;            // first staged column of the wave's fragments

    // ---- staging --------------------------------------------------------------
    // copy 0 of a chip row at its place, copy 1 shifted by one element
    auto stage_chip = [&](uint16_t *d0, uint4 v, uint32_t prev) {
        *reinterpret_cast<uint4 *>(d0) = v;
        *reinterpret_cast<uint4 *>(d0 + CS) =
            make_uint4((prev >> 16) | (v.x << 16), (v.x >> 16) | (v.y << 16),
                       (v.y >> 16) | (v.z << 16), (v.z >> 16) | (v.w << 16));
    };
    // chunk 0: target rows 32 (bb + 1) + o, o = 0 .. CH, of every u-block into
    // ring slot o; chip rows 0 .. CH - 1 into slot 0
    for (int k = threadIdx.x; k < NB * (CH + 1) * CPR; k += blockDim.x) {
        const int bb = k / ((CH + 1) * CPR), r = k - bb * (CH + 1) * CPR, o = r / CPR;
        const int q = r - o * CPR, gr = 32 * (bb + 1) + o, gc = 8 * q;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gr < WR && gc < W) v = *reinterpret_cast<const uint4 *>(src + (size_t)gr * W + gc);
        *reinterpret_cast<uint4 *>(ring + ((size_t)bb * RR + o) * TS + gc) = v;
    }
    for (int k = threadIdx.x; k < CH * CCH; k += blockDim.x) {
        const int i = k / CCH, q = k - i * CCH;
        const uint16_t *row = csrc + (size_t)i * CS;
        stage_chip(chip + (size_t)i * 2 * CS + 8 * q, *reinterpret_cast<const uint4 *>(row + 8 * q),
                   q ? *reinterpret_cast<const uint32_t *>(row + 8 * q - 2) : 0u);
    }
    // the refill of chunk c >= 1: RT rows -- target rows 32 (bb + 1) + CH c + 1
    // + rr (rr < CH) of every u-block, then chip rows CH c + i -- taken RPI rows
    // per wave-wide load (lane = 16-byte column chunk), row pairs round-robin
    // over the waves
    const int nw = __builtin_amdgcn_readfirstlane((int)blockDim.x >> 6);
    const int RT = NB * CH + CH;
    const int RPI = (CPR <= 32 && CCH <= 32) ? 2 : 1;
    const int rq = RPI == 2 ? (lane & 31) : lane;       // column chunk of this lane
    const int rh = RPI == 2 ? (lane >> 5) : 0;          // row of the pair
    const __amdgpu_buffer_rsrc_t trs =
        __builtin_amdgcn_make_buffer_rsrc((void *)src, (short)0, WR * W * 2, kBufferWord3);
    const __amdgpu_buffer_rsrc_t crs =
        __builtin_amdgcn_make_buffer_rsrc((void *)csrc, (short)0, CR * CS * 2, kBufferWord3);
    // ---- the wave's fragments at chip rows (0, 1), from global memory ---------
    // tile t, fragment f, lane (m, lk): target row 32 b + t + 2 m + d, columns
    // col0 + 16 f + 8 h .. + 7
    uint4 A0[NF], A1[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const int gc = col0 + 16 * f + 8 * h;
        const int r0 = 32 * b + 2 * lm + d;
        A0[f] = make_uint4(0u, 0u, 0u, 0u);
        A1[f] = make_uint4(0u, 0u, 0u, 0u);
        if (gc < W) {
            A0[f] = *reinterpret_cast<const uint4 *>(src + (size_t)r0 * W + gc);
            A1[f] = *reinterpret_cast<const uint4 *>(src + (size_t)(r0 + 1) * W + gc);
        }
    }
    floatx4 acc0[VT], acc1[VT], accE = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int v = 0; v < VT; ++v) acc0[v] = acc1[v] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();

    // this lane's chip fragment of K-step kappa: dwords 8 kappa + 4 h + 8 - (lm >> 1)
    // .. + 3 of copy (lm & 1) of the step's chip row d (a row is CS dwords: two copies)
    const uint32_t *cw = reinterpret_cast<const uint32_t *>(chip) + d * CS + (lm & 1) * (CS / 2) +
                         4 * h + 8 - (lm >> 1);
    // the rows entering lane 15 of each tile, and the packed row's fragments
    // (lanes m < VT of the last u-block's waves: fragment m + kappa)
    const bool in15 = lm == 15;
    const uint16_t *rb = ring + (size_t)b * RR * TS + col0 + 8 * h;
    // (lanes m >= VT load fragment VT - 1 + kappa as well: their M-rows are not kept)
    const uint16_t *eb = rb + 16 * min(lm, VT - 1);

    // the refill of the next chunk is spread over the steps: batch s of its
    // loads is issued before step s and stored after it (the slots it fills are
    // not read in this chunk)
    constexpr int NS = CH / 2;                     // steps per chunk
    constexpr int KB = refill_loads(NK);           // refill loads per thread and step
    // CS as an opaque value: the step's chip base is then computed once per
    // step, and each K-step's offset (32 kappa bytes) fits the loads' offset field
    int csr = CS;
    asm volatile("" : "+s"(csr));
    int base = 0;   // ring slot of offset CH c
    for (int c = 0; c < nchunks; ++c) {
        const bool more = c + 1 < nchunks;
        const uint32_t *slot = cw + (size_t)(c & 1) * CH * CS;
        int nbase = base + CH;                     // ring slot of offset CH (c + 1)
        nbase -= nbase >= RR ? RR : 0;
#pragma unroll 1
        for (int s = 0; s < NS; ++s) {
            // refill loads through buffer resources: an offset past the pair's
            // window or chip reads zeros, so no load needs a branch
            uint4 nv[KB];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const int p2 = (wave + (s * KB + j) * nw) * RPI;   // first row of the load (uniform)
                const int r = p2 + rh;
                int off = -1;                                   // byte offset; -1: none
                if (more && p2 < NB * CH) {
                    const int bb = r / CH, gr = 32 * (bb + 1) + CH * (c + 1) + 1 + (r - bb * CH);
                    if (rq < CPR && 8 * rq < W) off = 2 * (gr * W + 8 * rq);
                } else if (more && r < RT && rq < CCH) {
                    off = 2 * ((CH * (c + 1) + r - NB * CH) * CS + 8 * rq);
                }
                const bool tgt_row = p2 < NB * CH;
                nv[j] = __builtin_bit_cast(
                    uint4, __builtin_amdgcn_raw_buffer_load_b128(tgt_row ? trs : crs, off, 0, 0));
            }
            // the step's chip rows 2 s, 2 s + 1 of the slot; the rows entering
            // lane 15 of tile t: ring offset CH c + 2 s + t + d
            const uint32_t *cr = slot + (size_t)(2 * s) * csr;
            int s0 = base + 2 * s + d, s1 = s0 + 1;
            s0 -= s0 >= RR ? RR : 0;
            s1 -= s1 >= RR ? RR : 0;
            const uint16_t *x0p = rb + (size_t)s0 * TS, *x1p = rb + (size_t)s1 * TS;
            const uint16_t *e0p = eb + (size_t)s0 * TS;
            auto ldB = [&](int kap) {
                return make_uint4(cr[8 * kap], cr[8 * kap + 1], cr[8 * kap + 2], cr[8 * kap + 3]);
            };
            // the step's K-steps; EG: this wave holds the packed row too
            auto body = [&](auto edge_tag) {
                constexpr bool EG = decltype(edge_tag)::value;
                uint4 B[NK];
                B[0] = ldB(0);
#pragma unroll
                for (int kap = 0; kap < NF; ++kap) {
                    if (kap + 1 < NK) B[kap + 1] = ldB(kap + 1);
                    if (kap < NK) {
                        // the packed row's fragment, loaded at the start of the
                        // K-step; its MFMA goes last
                        uint4 E;
                        if (edge) E = *reinterpret_cast<const uint4 *>(e0p + 16 * kap);
#pragma unroll
                        for (int v = 0; v < VT; ++v) {
                            acc0[v] = mfma32<FP16>(A0[v + kap], B[kap], acc0[v]);
                            acc1[v] = mfma32<FP16>(A1[v + kap], B[kap], acc1[v]);
                        }
                        if (edge) accE = mfma32<FP16>(E, B[kap], accE);
                    }
                    // fragments that have served their last K-step of this step,
                    // kShift at a time: lanes 0..14 take the next M-row's target
                    // row, lane 15 the new row
                    if (kap % kShift == kShift - 1 || kap == NF - 1) {
                        const int f0 = kap - kap % kShift;
#pragma unroll
                        for (int f = f0; f <= kap; ++f) {
                            A0[f] = shift_up1(A0[f], A0[f]);
                            A1[f] = shift_up1(A1[f], A1[f]);
                        }
                        if (in15) {
#pragma unroll
                            for (int f = f0; f <= kap; ++f) {
                                A0[f] = *reinterpret_cast<const uint4 *>(x0p + 16 * f);
                                A1[f] = *reinterpret_cast<const uint4 *>(x1p + 16 * f);
                            }
                        }
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            };
            body(std::true_type{});
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const int p2 = (wave + (s * KB + j) * nw) * RPI;
                const int r = p2 + rh;
                if (more && p2 < NB * CH) {
                    const int bb = r / CH;
                    int sl = nbase + 1 + (r - bb * CH);
                    sl -= sl >= RR ? RR : 0;
                    if (rq < CPR)
                        *reinterpret_cast<uint4 *>(ring + ((size_t)bb * RR + sl) * TS + 8 * rq) =
                            nv[j];
                } else if (more && p2 < RT) {
                    // the element before this chunk: the last one of the previous
                    // lane's chunk (zero at the start of a row)
                    const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(
                        0, (int)nv[j].w, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
                    if (r < RT && rq < CCH)
                        stage_chip(chip + ((size_t)((c + 1) & 1) * CH + r - NB * CH) * 2 * CS +
                                       8 * rq,
                                   nv[j], rq ? prev : 0u);
                }
            }
        }
        __syncthreads();
        base = nbase;
    }

    // ---- outputs: lane (n, lk) holds M-rows 4 lk + e of column n; with
    // accumulate set they are added to what the surface holds ----------------
    float *dst = partial + ((size_t)pair * nsplit + ks) * P * P;
    auto put = [&](size_t i, float x) { dst[i] = accumulate ? dst[i] + x : x; };
#pragma unroll
    for (int v = 0; v < VT; ++v) {
        const int vv = col0 + 16 * v + lm;
        if (vv >= P) continue;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int u0 = 32 * b + 2 * (4 * lk + e);
            put((size_t)u0 * P + vv, acc0[v][e]);
            put((size_t)(u0 + 1) * P + vv, acc1[v][e]);
        }
    }
    if (edge) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int m = 4 * lk + e, vv = col0 + 16 * m + lm;
            if (m < VT && vv < P) put((size_t)(P - 1) * P + vv, accE[e]);
        }
    }
}

}  // namespace packed_detail

// LDS bytes of one workgroup: the target ring of every u-block and two chip slots
inline size_t packed_lds_bytes(int nk, int vt, int ch, int nb, int nvg)
{
    const size_t ring = (size_t)nb * (2 * ch + 1) * packed_detail::packed_row_stride(nk, vt, nvg);
    return (ring + (size_t)2 * ch * 2 * (16 * nk + 16)) * sizeof(uint16_t);
}

// whether one chunk's refill (the target rows of every u-block and the chip
// rows, one or two rows per wave-wide load) fits the refill loads per thread and step
inline bool packed_refill_fits(int nk, int vt, int ch, int nb, int nvg)
{
    const int cpr = (packed_detail::packed_row_stride(nk, vt, nvg) - 8) / 8, cch = 2 * nk + 2;
    if (cpr > 64 || cch > 64) return false;
    const int rpi = (cpr <= 32 && cch <= 32) ? 2 : 1;
    const int pairs = (nb * ch + ch + rpi - 1) / rpi, waves = nb * nvg;
    return (pairs + waves - 1) / waves <= packed_detail::refill_loads(nk) * (ch / 2);
}

// the K-step counts built: 2 .. 11 (chips of 16 .. 160 samples) and 17 (256, the
// refined chips of 128-sample chips)
__host__ __device__ constexpr bool packed_nk_built(int nk)
{
    return (nk >= 2 && nk <= 11) || nk == 17;
}

// VGPRs of a wave: two tiles' fragments (8 per fragment), accumulators (8 per
// v-tile), about 40 for the chip fragments, the packed row's loads and
// addresses, and 4 per refill load past two (fitted to the compiled kernels);
// a (K-steps, v-tiles, waves per SIMD) kind is built only where they fit the
// register file at that occupancy (168 at three waves, 256 at two)
__host__ __device__ constexpr int packed_vgprs(int nk, int vt)
{
    return 16 * vt + 8 * nk + 40 + 4 * (packed_detail::refill_loads(nk) - 2);
}
__host__ __device__ constexpr bool packed_kind_fits(int nk, int vt, int wpe)
{
    return packed_vgprs(nk, vt) <= (wpe == 3 ? 168 : 256);
}

// split shapes (hi / lo operands) run three passes over the same tiling:
// hi x hi written, then hi x lo and lo x hi added
template <int NK>
hipError_t launch_correlate_packed_nk(hipStream_t s, const MixedShape &sh, const uint16_t *chip_q,
                                      const uint16_t *tgt_q, int pairs, int C, int W, bool fp16,
                                      float *partial, const uint16_t *chip_lo,
                                      const uint16_t *tgt_lo)
#ifdef AMPCOR_MIXED_INSTANTIATE
{
    using namespace packed_detail;
    const int P = W - C + 1;
    const dim3 block(64 * sh.wg_ut);
    const int passes = sh.split ? 3 : 1;
    // one workgroup per pair and chip-row range: at most 2^31 - 1 of them per launch
    const int ns = sh.nsplit;
    const int chunk = std::min(pairs, 0x7fffffff / ns);
    for (int pass = 0; pass < passes; ++pass)
    for (int first = 0; first < pairs; first += chunk) {
        const int n = std::min(chunk, pairs - first);
        const uint16_t *cq = (pass == 2 ? chip_lo : chip_q) + (size_t)first * C * sh.cs;
        const uint16_t *tq = (pass == 1 ? tgt_lo : tgt_q) + (size_t)first * W * W;
        float *pt = partial + (size_t)first * ns * P * P;
        const int acc = pass > 0;
        bool launched = false;
        auto kind = [&](auto vt, auto ch, auto wpe) {
            constexpr int VTV = decltype(vt)::value, CHV = decltype(ch)::value;
            constexpr int WPEV = decltype(wpe)::value;
            if constexpr (packed_kind_fits(NK, VTV, WPEV)) {
                if (launched || sh.wg_vt != VTV || sh.packed != CHV || sh.wpe != WPEV) return;
                launched = true;
                if (fp16)
                    hipLaunchKernelGGL((k_correlate_packed<NK, VTV, CHV, WPEV, true>),
                                       dim3(n * ns), block, sh.lds_bytes, s, cq, tq, C, W, P,
                                       sh.tiles_v, ns, pt, acc);
                else
                    hipLaunchKernelGGL((k_correlate_packed<NK, VTV, CHV, WPEV, false>),
                                       dim3(n * ns), block, sh.lds_bytes, s, cq, tq, C, W, P,
                                       sh.tiles_v, ns, pt, acc);
            }
        };
        using I2 = std::integral_constant<int, 2>;
        using I3 = std::integral_constant<int, 3>;
        using I4 = std::integral_constant<int, 4>;
        using I5 = std::integral_constant<int, 5>;
        using I8 = std::integral_constant<int, 8>;
        kind(I2{}, I4{}, I3{});
        kind(I3{}, I4{}, I3{});
        kind(I3{}, I4{}, I2{});
        kind(I5{}, I4{}, I2{});
        kind(I2{}, I8{}, I3{});
        kind(I3{}, I8{}, I3{});
        kind(I3{}, I8{}, I2{});
        kind(I5{}, I8{}, I3{});
        kind(I5{}, I8{}, I2{});
        if (!launched) return hipErrorInvalidValue;
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
#else
    ;
#endif

}  // namespace ampcor_hip
