// fp8.hip -- tiling and launch of the FP8 coarse contraction (fp8_correlate.h).
#include "fp8_correlate.h"

#include <algorithm>

namespace ampcor_hip {

Fp8Shape choose_fp8_shape(int C, int P, int nsplit, int max_vt)
{
    Fp8Shape s{};
    s.nk = (C + 15 + 15) / 16;
    s.cs = 16 * s.nk + 16;
    const int tiles = (P + 15) / 16;
    const int vmax = max_vt > 0 && max_vt < fp8_detail::kMaxVT ? max_vt : fp8_detail::kMaxVT;
    const int vblocks = (tiles + vmax - 1) / vmax;
    s.wg_vt = (tiles + vblocks - 1) / vblocks;
    const int ublocks = (tiles + fp8_detail::kMaxWaves - 1) / fp8_detail::kMaxWaves;
    s.wg_ut = (tiles + ublocks - 1) / ublocks;   // waves, spread evenly over the blocks
    // chip rows in steps of eight: ranges of a multiple of 8 rows (zero past C)
    const int c8 = (C + 7) / 8 * 8;
    auto can_split = [&](int ns) { return c8 % (8 * ns) == 0; };
    s.nsplit = 1;
    const size_t budget = 64 * 1024;
    while (fp8_lds_bytes(s.nk, s.wg_vt, s.wg_ut, c8 / s.nsplit) > budget) {
        if (can_split(2 * s.nsplit) && c8 / (2 * s.nsplit) >= 16) s.nsplit *= 2;
        else if (s.wg_ut > 1) s.wg_ut = (s.wg_ut + 1) / 2;
        else break;
    }
    // a requested split (tuning)
    if (nsplit > s.nsplit && (nsplit & (nsplit - 1)) == 0 && can_split(nsplit) && c8 / nsplit >= 8)
        s.nsplit = nsplit;
    s.ilen = c8 / s.nsplit;
    s.tiles_u = (tiles + s.wg_ut - 1) / s.wg_ut;
    s.tiles_v = (tiles + s.wg_vt - 1) / s.wg_vt;
    s.lds_bytes = fp8_lds_bytes(s.nk, s.wg_vt, s.wg_ut, s.ilen);
    return s;
}

namespace {

template <int NK>
hipError_t launch_fp8_nk(hipStream_t s, const Fp8Shape &sh, const uint8_t *chip_q,
                         const uint8_t *tgt_q, int pairs, int C, int W, float *partial)
{
    using namespace fp8_detail;
    const int P = W - C + 1;
    const size_t per = (size_t)sh.tiles_u * sh.tiles_v * sh.nsplit;   // workgroups per pair
    const dim3 block(64 * sh.wg_ut);
    // a launch's work-items fit in 32 bits: pairs go in chunks
    const int chunk = std::min(pairs, launch_pair_cap(per * block.x));
#define AMPCOR_FP8_VT(VTV)                                                                         \
    case VTV:                                                                                      \
        hipLaunchKernelGGL((k_correlate_fp8<NK, VTV>), grid, block, sh.lds_bytes, s, cq, tq, C, W, \
                           P, sh.ilen, sh.wg_ut, sh.tiles_u, sh.tiles_v, sh.nsplit, pt);          \
        break
    for (int first = 0; first < pairs; first += chunk) {
        const int n = std::min(chunk, pairs - first);
        const dim3 grid((unsigned)(n * per));
        const uint8_t *cq = chip_q + (size_t)first * C * sh.cs;
        const uint8_t *tq = tgt_q + (size_t)first * W * W;
        float *pt = partial + (size_t)first * sh.nsplit * P * P;
        switch (sh.wg_vt) {
            AMPCOR_FP8_VT(1);
            AMPCOR_FP8_VT(2);
            AMPCOR_FP8_VT(3);
            AMPCOR_FP8_VT(4);
            AMPCOR_FP8_VT(5);
        default: return hipErrorInvalidValue;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
#undef AMPCOR_FP8_VT
    return hipSuccess;
}

}  // namespace

hipError_t launch_correlate_fp8(hipStream_t s, const Fp8Shape &sh, const uint8_t *chip_q,
                                const uint8_t *tgt_q, int pairs, int C, int W, float *partial)
{
    const int P = W - C + 1;
    if (sh.nk != (C + 15 + 15) / 16 || sh.cs != 16 * sh.nk + 16 || sh.ilen % 8 ||
        sh.ilen * sh.nsplit < C || sh.wg_ut < 1 || sh.wg_ut > fp8_detail::kMaxWaves ||
        sh.tiles_u * sh.wg_ut * 16 < P || sh.tiles_v * sh.wg_vt * 16 < P)
        return hipErrorInvalidValue;
#define AMPCOR_FP8_NK(N)                                                                           \
    case N: return launch_fp8_nk<N>(s, sh, chip_q, tgt_q, pairs, C, W, partial)
    switch (sh.nk) {
        AMPCOR_FP8_NK(2); AMPCOR_FP8_NK(3); AMPCOR_FP8_NK(4); AMPCOR_FP8_NK(5);
        AMPCOR_FP8_NK(6); AMPCOR_FP8_NK(7); AMPCOR_FP8_NK(8); AMPCOR_FP8_NK(9);
    default: return hipErrorInvalidValue;
    }
#undef AMPCOR_FP8_NK
}

}  // namespace ampcor_hip
