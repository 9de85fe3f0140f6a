// plan.hip -- the C ABI (include/ampcor_hip.h): rasters, plans, adjust.
//
// A plan mirrors the reference's Sequential worker
// (lib/libampcor_cuda/correlators/Sequential.{h,icc}) without its host arena:
// pairs are registered as descriptors (raster, upper-left corner) and adjust
// gathers the tiles straight from device-resident rasters.  All device
// buffers are allocated once, at plan creation, and every stage runs on the
// plan's own stream with no intermediate synchronization.
#include "plan_internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace ampcor_hip;

namespace {
thread_local std::string g_error;
}  // namespace

namespace ampcor_hip {

int fail(int code, const std::string &message)
{
    g_error = message;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return fail(AMPCOR_HIP_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace ampcor_hip

namespace {

// one entry of the band-limited upsampling kernel (closed form):
// t = p/rf - n; odd N: sin(pi t) / (N sin(pi t / N));
// even N (Nyquist bin split): sin(pi t) cot(pi t / N) / N
double upsample_kernel(int N, int rf, int p, int n)
{
    if (p % rf == 0) return (p / rf == n) ? 1.0 : 0.0;
    const double t = (double)(p - rf * n) / rf;
    const double a = M_PI * t;
    if (N % 2) return std::sin(a) / (N * std::sin(a / N));
    return std::sin(a) * std::cos(a / N) / (N * std::sin(a / N));
}

// periodic Dirichlet kernel of an odd-size R grid at x = d / zf
double dirichlet(int R, int zf, int d)
{
    if (d % zf == 0) return ((d / zf) % R == 0) ? 1.0 : 0.0;
    const double x = (double)d / zf;
    return std::sin(M_PI * x) / (R * std::sin(M_PI * x / R));
}

}  // namespace

namespace {

constexpr int kTuneMaxSplitCoarse = 4;
constexpr int kTuneMaxSplitRefined = 16;

int check_device(int device)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(AMPCOR_HIP_EDEVICE, "no HIP device available");
    if (device < 0 || device >= count)
        return fail(AMPCOR_HIP_EINVAL, "device index " + std::to_string(device) + " out of range");
    return AMPCOR_HIP_OK;
}

// host tables bound for device buffers, copied together once the plan's
// buffers are laid out
struct Uploads {
    std::vector<std::pair<void *, std::vector<unsigned char>>> items;
    template <typename T>
    int put(DeviceBuffer<T> &buf, const std::vector<T> &host, Carver &cv, const char *what)
    {
        const hipError_t e = buf.carve(cv, host.size());
        if (e != hipSuccess)
            return fail(AMPCOR_HIP_ENOMEM, std::string(what) + ": " + hipGetErrorString(e));
        const auto *b = reinterpret_cast<const unsigned char *>(host.data());
        items.emplace_back(buf.ptr, std::vector<unsigned char>(b, b + host.size() * sizeof(T)));
        return AMPCOR_HIP_OK;
    }
    int flush(hipStream_t s)
    {
        for (auto &it : items)
            HIP_TRY(hipMemcpyAsync(it.first, it.second.data(), it.second.size(),
                                   hipMemcpyHostToDevice, s),
                    "uploading plan tables");
        HIP_TRY(hipStreamSynchronize(s), "uploading plan tables");
        return AMPCOR_HIP_OK;
    }
};

// the upsampling matrix U[p][n] = kernel(N, rf, p, n), padded with zeros to
// (rf N rounded up to 16) x (N rounded up to 16) for the matrix-core kernel
int upload_upsampler(DeviceBuffer<float> &U, DeviceBuffer<float> &g, int N, int rf, Carver &cv,
                     Uploads &up)
{
    const int M = rf * N;
    const int Np = (N + 15) / 16 * 16, Mp = (M + 15) / 16 * 16;
    std::vector<float> u((size_t)Mp * Np, 0.0f);
    for (int p = 0; p < M; ++p)
        for (int n = 0; n < N; ++n) u[(size_t)p * Np + n] = (float)upsample_kernel(N, rf, p, n);
    int rc = up.put(U, u, cv, "allocating an upsampling matrix");
    if (rc || rf != 2) return rc;
    // the half-sample shift is circulant: U[2a + 1][n] = g((a - n) mod N)
    std::vector<float> row(N);
    for (int d = 0; d < N; ++d) row[d] = (float)upsample_kernel(N, rf, 2 * d + 1, 0);
    return up.put(g, row, cv, "allocating an upsampling matrix");
}

// the half-sample circulant H[a][n] = g((a - n) mod N) as FP16 fragments of the
// 16 x 16 x 32 MFMA: block (mt, kc), lane l, value e holds
// H[16 mt + l % 16][32 kc + 8 (l / 16) + e], zero outside N x N
int upload_fragments(DeviceBuffer<uint16_t> &hf, int N, Carver &cv, Uploads &up)
{
    const int Np = (N + 31) / 32 * 32, MT = Np / 16, KC = Np / 32;
    std::vector<double> g(N);
    for (int d = 0; d < N; ++d) g[d] = upsample_kernel(N, 2, 2 * d + 1, 0);
    std::vector<uint16_t> h((size_t)upsample_fragment_count(N), 0);
    size_t at = 0;
    for (int mt = 0; mt < MT; ++mt)
        for (int kc = 0; kc < KC; ++kc)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e, ++at) {
                    const int row = 16 * mt + l % 16, col = 32 * kc + 8 * (l / 16) + e;
                    if (row < N && col < N) {
                        const _Float16 v = (_Float16)g[((row - col) % N + N) % N];
                        h[at] = __builtin_bit_cast(uint16_t, v);
                    }
                }
    return up.put(hf, h, cv, "allocating the upsampling fragments");
}

// the zoom's FP16 screen: the Dirichlet matrix M[p][a] = D(p - zf a) as
// 16 x 32 MFMA fragments (zoom_fragment_count, in both K orders), and the
// bounds on the screen's error.  With Lambda = max_p sum_a |D(p - zf a)| (the Lebesgue constant of
// the zoom), every |H[p][b]| <= Lambda max|g| and sum_b |D(q - zf b)| <=
// Lambda.  Rounding g, D, D again and H to FP16 (relative 2^-11 each, four
// times) and FP32 accumulation over KP = R padded to 32 terms in each product
// give |f_screen - f| <= Lambda^2 max|g| (4 2^-11 + 2 KP 2^-24); FP16
// subnormals (absolute 2^-25 per rounding) add 2^-25 (Lambda^2 + 3 KP Lambda).
// After the first product alone: Lambda max|g| (2 2^-11 + KP 2^-24) and
// 2^-25 (Lambda + KP).  All carry a 1.25 margin for the second-order terms.
int upload_zoom_screen(ampcor_hip_plan *p, const std::vector<double> &table /* d + Z */,
                       Carver &cv, Uploads &up)
{
    const int R = p->R, zf = p->zf, Z = p->Z;
    const int ZT = (Z + 15) / 16, KC = (R + 31) / 32, KP = 32 * KC;
    std::vector<uint16_t> h((size_t)zoom_fragment_count(R, zf), 0), hp(h.size(), 0);
    size_t at = 0;
    for (int zt = 0; zt < ZT; ++zt)
        for (int kc = 0; kc < KC; ++kc)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e, ++at) {
                    const int row = 16 * zt + l % 16, lk = l / 16;
                    const int a = 32 * kc + 8 * lk + e;
                    const int b = 32 * kc + (e < 4 ? 4 * lk + e : 16 + 4 * lk + e - 4);
                    if (row < Z && a < R)
                        h[at] = __builtin_bit_cast(uint16_t, (_Float16)table[row - zf * a + Z]);
                    if (row < Z && b < R)
                        hp[at] = __builtin_bit_cast(uint16_t, (_Float16)table[row - zf * b + Z]);
                }
    double lambda = 0.0;
    for (int row = 0; row < Z; ++row) {
        double sum = 0.0;
        for (int a = 0; a < R; ++a) sum += std::fabs(table[row - zf * a + Z]);
        lambda = std::max(lambda, sum);
    }
    auto &b = p->zoom_bounds;
    b.ecoef = 1.25 * lambda * lambda * (4.0 * std::ldexp(1.0, -11) + 2.0 * KP * std::ldexp(1.0, -24));
    b.eabs = 1.25 * std::ldexp(1.0, -25) * (lambda * lambda + 3.0 * KP * lambda);
    // after the first product alone (g and D rounded, FP32 accumulation)
    b.hcoef = (float)(1.25 * lambda * (2.0 * std::ldexp(1.0, -11) + KP * std::ldexp(1.0, -24)));
    b.habs = (float)(1.25 * std::ldexp(1.0, -25) * (lambda + KP));
    b.lambda = (float)(lambda * (1.0 + std::ldexp(1.0, -20)));
    if (int rc = up.put(p->zfrag, h, cv, "allocating the zoom screen")) return rc;
    if (int rc = up.put(p->zfragp, hp, cv, "allocating the zoom screen")) return rc;
    const hipError_t e = p->zstats.carve(cv, 3);
    return e == hipSuccess ? AMPCOR_HIP_OK
                           : fail(AMPCOR_HIP_ENOMEM, std::string("allocating the zoom screen: ") +
                                                        hipGetErrorString(e));
}

int register_tile(ampcor_hip_plan *plan, const ampcor_hip_raster *raster, size_t idx,
                  size_t bl, size_t bs, size_t el, size_t es, bool reference)
{
    if (!plan || !raster) return fail(AMPCOR_HIP_EINVAL, "null plan or raster");
    if (idx >= plan->pairs)
        return fail(AMPCOR_HIP_EINVAL, "pair index " + std::to_string(idx) + " out of range [0, " +
                                           std::to_string(plan->pairs) + ")");
    if (raster->device != plan->device)
        return fail(AMPCOR_HIP_EINVAL, "raster lives on another device than the plan");
    if (plan->dtype >= 0 && raster->dtype != plan->dtype)
        return fail(AMPCOR_HIP_EINVAL, "all rasters of a plan must share one pixel type");
    const size_t edge = reference ? plan->C : plan->W;
    if (el < bl || es < bs || el - bl != edge || es - bs != edge)
        return fail(AMPCOR_HIP_EINVAL, std::string(reference ? "reference" : "target") +
                                           " slice shape does not match the plan's " +
                                           (reference ? "chip" : "window"));
    if (el > raster->lines || es > raster->samples)
        return fail(AMPCOR_HIP_EINVAL, "slice extends beyond the raster");
    if (raster->samples > 0x7fffffff || raster->lines > 0x7fffffff)
        return fail(AMPCOR_HIP_EINVAL, "raster too large");
    plan->dtype = raster->dtype;
    PairDesc &d = plan->desc[idx];
    if (reference) {
        d.ref = raster->data;
        d.ref_stride = (int32_t)raster->samples;
        d.ref_line = (int32_t)bl;
        d.ref_sample = (int32_t)bs;
        plan->have_ref[idx] = 1;
    } else {
        d.tgt = raster->data;
        d.tgt_stride = (int32_t)raster->samples;
        d.tgt_line = (int32_t)bl;
        d.tgt_sample = (int32_t)bs;
        plan->have_tgt[idx] = 1;
    }
    plan->desc_dirty = true;
    plan->adjusted = false;
    return AMPCOR_HIP_OK;
}

}  // namespace

namespace ampcor_hip {

int plan_configure(ampcor_hip_plan *p, int device, size_t pairs, size_t ref_lines,
                   size_t ref_samples, size_t tgt_lines, size_t tgt_samples, size_t refine_margin,
                   size_t refine_factor, size_t zoom_factor, int precision)
{
    if (pairs == 0) return fail(AMPCOR_HIP_EINVAL, "a plan needs at least one pair");
    // launches with one workgroup (at most 1024 threads) per pair stay within
    // 32-bit work-item counts; those with more workgroups per pair go in chunks
    if (pairs > kMaxPlanPairs)
        return fail(AMPCOR_HIP_EINVAL, "at most " + std::to_string(kMaxPlanPairs) +
                                           " pairs per plan: split the batch");
    // the kernels assume square tiles (Sequential.icc:56-60 makes the same assumption)
    if (ref_lines != ref_samples || tgt_lines != tgt_samples)
        return fail(AMPCOR_HIP_EINVAL, "chips and windows must be square");
    const size_t C = ref_lines, W = tgt_lines;
    if (C < 2 || C > 512) return fail(AMPCOR_HIP_EINVAL, "chip edge must be in [2, 512]");
    if (W < C || (W - C) % 2) return fail(AMPCOR_HIP_EINVAL, "window - chip must be even and >= 0");
    if (W > 512) return fail(AMPCOR_HIP_EINVAL, "window edge must be <= 512");
    const size_t m = (W - C) / 2;
    if (refine_margin > m)
        return fail(AMPCOR_HIP_EINVAL, "refine margin exceeds the search margin");
    if (refine_factor < 1 || refine_factor > 8) return fail(AMPCOR_HIP_EINVAL, "refine factor must be in [1, 8]");
    if (zoom_factor < 1 || zoom_factor > 31) return fail(AMPCOR_HIP_EINVAL, "zoom factor must be in [1, 31]");
    const size_t E = C + 2 * refine_margin, Cr = refine_factor * C, Wr = refine_factor * E;
    const size_t R = Wr - Cr + 1;
    if (Wr > 512) return fail(AMPCOR_HIP_EINVAL, "refined window edge must be <= 512");
    if (R > 129) return fail(AMPCOR_HIP_EINVAL, "refined surface too large (2 rf rm + 1 > 129)");
    if (!zoom_fits((int)R, (int)zoom_factor))
        return fail(AMPCOR_HIP_EINVAL, "zoom factor too large for this refined surface (the zoom "
                                       "of one pair must fit a workgroup's 160 KiB of LDS)");
    if (precision < AMPCOR_HIP_FP32 || precision > AMPCOR_HIP_FP8)
        return fail(AMPCOR_HIP_EINVAL, "unknown precision (FP32, BF16, FP16, BF16X3, FP16X3 or FP8)");
    if (precision != AMPCOR_HIP_FP32 && (C % 16 || C * refine_factor > 256))
        return fail(AMPCOR_HIP_EINVAL, "the mixed forms need a chip edge that is a multiple of 16 "
                                       "and a refined chip edge <= 256");
    // the FP8 coarse stage runs as one gather-quantize pass and the fused
    // normalizer (its window sums come from the FP8 operand)
    if (precision == AMPCOR_HIP_FP8 &&
        (!gather_prep_fits((int)C, (int)W) || !normalize_fuses((int)W, (int)C)))
        return fail(AMPCOR_HIP_EINVAL, "the FP8 form needs a chip edge <= 128 and (window - chip + 1) <= 144");
    if (int rc = check_device(device)) return rc;

    p->device = device;
    p->precision = precision;
    p->pairs = pairs;
    p->C = (int)C;
    p->W = (int)W;
    p->m = (int)m;
    p->P = (int)(W - C + 1);
    p->rf = (int)refine_factor;
    p->rm = (int)refine_margin;
    p->E = (int)E;
    p->Cr = (int)Cr;
    p->Wr = (int)Wr;
    p->R = (int)R;
    p->zf = (int)zoom_factor;
    p->Z = p->zf * p->R;
    p->desc.assign(pairs, PairDesc{});
    p->have_ref.assign(pairs, 0);
    p->have_tgt.assign(pairs, 0);
    {
        // tiling overrides of the FP32 contractions: AMPCOR_HIP_TILE_COARSE="bu[,jblock]",
        // AMPCOR_HIP_TILE_REFINED="bu,nsplit[,jblock]"
        int cbu = 0, cjb = 0, rbu = 0, rns = 0, rjb = 0;
        if (const char *e = std::getenv("AMPCOR_HIP_TILE_COARSE")) std::sscanf(e, "%d,%d", &cbu, &cjb);
        if (const char *e = std::getenv("AMPCOR_HIP_TILE_REFINED"))
            std::sscanf(e, "%d,%d,%d", &rbu, &rns, &rjb);
        p->coarse_shape = choose_correlate_shape(p->C, p->P, true, cbu, 0, cjb);
        p->refined_shape = choose_correlate_shape(p->Cr, p->R, false, rbu, rns, rjb);
    }
    p->mixed = precision != AMPCOR_HIP_FP32;
    p->split = precision == AMPCOR_HIP_BF16X3 || precision == AMPCOR_HIP_FP16X3;
    p->fp8 = precision == AMPCOR_HIP_FP8;
    p->fp16 = precision == AMPCOR_HIP_FP16 || precision == AMPCOR_HIP_FP16X3 || p->fp8;
    if (p->mixed) {
        p->mixed_coarse = choose_mixed_shape(p->C, p->P);
        p->mixed_refined = choose_mixed_shape(p->Cr, p->R, 0, p->split);
        // the packed form where the surface has 32 NB + 1 placements per axis
        // (no padded placement rows; split operands in three passes), else the
        // refined contraction streams its chip rows where it can (plain
        // operands; one partial surface per pair); the autotuner weighs every form
        // (AMPCOR_HIP_PACKED_SPLIT="coarse,refined" sets its chip-row ranges)
        int pcs = 1, prs = 1;
        if (const char *e = std::getenv("AMPCOR_HIP_PACKED_SPLIT")) std::sscanf(e, "%d,%d", &pcs, &prs);
        MixedShape pc = choose_mixed_packed_shape(p->C, p->P, 0, 0, false, pcs);
        if (!pc.nk) pc = choose_mixed_packed_shape(p->C, p->P);
        if (pc.nk) p->mixed_coarse = pc;
        MixedShape pr = choose_mixed_packed_shape(p->Cr, p->R, 0, 0, p->split, prs);
        if (!pr.nk) pr = choose_mixed_packed_shape(p->Cr, p->R, 0, 0, p->split);
        const MixedShape st = p->split ? MixedShape{} : choose_mixed_stream_shape(p->Cr, p->R);
        if (pr.nk) p->mixed_refined = pr;
        else if (st.nk) p->mixed_refined = st;
    }
    if (p->fp8) p->fp8_coarse = choose_fp8_shape(p->C, p->P);
    return AMPCOR_HIP_OK;
}

int plan_build(ampcor_hip_plan *p, Carver &cv, hipStream_t s, bool upload)
{
    const size_t N = p->pairs;
    const size_t C = p->C, W = p->W, Cr = p->Cr, Wr = p->Wr, R = p->R;
    Uploads up;
    struct {
        hipError_t e;
    } r{hipSuccess};
    auto A = [&](hipError_t x) {
        if (r.e == hipSuccess) r.e = x;
    };
    A(p->d_desc.carve(cv, N));
    A(p->amp_ref.carve(cv, N * C * C));
    A(p->amp_tgt.carve(cv, N * W * W));
    A(p->rT.carve(cv, N * C * C));
    A(p->stats.carve(cv, N));
    A(p->sums.carve(cv, N * p->P * p->P));
    A(p->sums_r.carve(cv, N * R * R));
    // partial buffers leave room for the autotuner's largest candidate splits
    const size_t ns_c = std::max({p->coarse_shape.nsplit, p->mixed ? p->mixed_coarse.nsplit : 1,
                                  p->mixed ? kTuneMaxSplitCoarse : 1,
                                  p->fp8 ? p->fp8_coarse.nsplit : 1});
    const size_t ns_r = std::max({p->refined_shape.nsplit, p->mixed ? p->mixed_refined.nsplit : 1,
                                  kTuneMaxSplitRefined});
    A(p->partial.carve(cv, N * ns_c * p->P * p->P));
    A(p->gamma.carve(cv, N * p->P * p->P));
    A(p->loc.carve(cv, N));
    A(p->window.carve(cv, N));
    A(p->peak.carve(cv, N));
    A(p->snr.carve(cv, N));
    A(p->amp_ref_r.carve(cv, N * Cr * Cr));
    A(p->amp_tgt_r.carve(cv, N * Wr * Wr));
    A(p->rT_r.carve(cv, N * Cr * Cr));
    A(p->stats_r.carve(cv, N));
    A(p->partial_r.carve(cv, N * ns_r * R * R));
    if (p->mixed) {
        // (the FP8 operands, one byte per value, fit the same buffers)
        A(p->chip_q.carve(cv, N * C * p->mixed_coarse.cs));
        A(p->tgt_q.carve(cv, N * W * W));
        A(p->chip_q_r.carve(cv, N * Cr * p->mixed_refined.cs));
        A(p->tgt_q_r.carve(cv, N * Wr * Wr));
        if (p->split) {
            A(p->chip_lo_r.carve(cv, N * Cr * p->mixed_refined.cs));
            A(p->tgt_lo_r.carve(cv, N * Wr * Wr));
        }
        A(p->mstats.carve(cv, N));
        A(p->mstats_r.carve(cv, N));
    }
    if (p->fp8) {
        A(p->guard.carve(cv, 3));
        A(p->near.carve(cv, N));
    }
    A(p->gamma_r.carve(cv, N * R * R));
    A(p->offsets.carve(cv, 2 * N));
    A(p->peak_r.carve(cv, N));
    A(p->zoom_loc.carve(cv, N));
    A(p->zthr.carve(cv, N));
    A(p->zwin.carve(cv, 4 * N));
    A(p->zflag.carve(cv, N));
    A(p->zncand.carve(cv, N));
    A(p->zcand.carve(cv, N * kZoomList));
    if (r.e != hipSuccess) {
        std::string msg = std::string("allocating plan buffers for ") + std::to_string(N) +
                          " pairs: " + hipGetErrorString(r.e);
        return fail(AMPCOR_HIP_ENOMEM, msg);
    }
    if (int rc = upload_upsampler(p->U_ref, p->g_ref, p->C, p->rf, cv, up)) return rc;
    if (int rc = upload_upsampler(p->U_tgt, p->g_tgt, p->E, p->rf, cv, up)) return rc;
    // the BF16/FP16/FP8 forms upsample tiles up to 96 samples (160 for real
    // rasters) on FP16 matrix cores (the split forms keep FP32)
    if (p->mixed && !p->split && p->rf == 2) {
        if (upsample_uses_fragments(p->C))
            if (int rc = upload_fragments(p->hfrag_ref, p->C, cv, up)) return rc;
        if (upsample_uses_fragments(p->E))
            if (int rc = upload_fragments(p->hfrag_tgt, p->E, cv, up)) return rc;
    }
    // larger tiles upsample in two banded passes through scratch, shared by
    // the two roles (they run one after the other on the plan stream)
    {
        int nmax = 0;
        if (upsample_needs_scratch(p->C, p->rf)) nmax = std::max(nmax, p->C);
        if (upsample_needs_scratch(p->E, p->rf)) nmax = std::max(nmax, p->E);
        if (nmax) {
            const size_t nb = (nmax + 15) / 16;
            hipError_t ea = p->up_T.carve(cv, N * (size_t)nmax * nmax);
            if (ea == hipSuccess) ea = p->up_band.carve(cv, N * nb);
            if (ea == hipSuccess) ea = p->up_rows.carve(cv, N * nb);
            if (ea != hipSuccess)
                return fail(AMPCOR_HIP_ENOMEM, std::string("allocating the upsampling scratch: ") +
                                                   hipGetErrorString(ea));
            p->up_scratch = UpsampleScratch{p->up_T.ptr, p->up_band.ptr, p->up_rows.ptr};
        }
    }
    {
        std::vector<double> d(2 * p->Z + 1);
        for (int k = -p->Z; k <= p->Z; ++k) d[k + p->Z] = dirichlet(p->R, p->zf, k);
        if (int rc = up.put(p->D, d, cv, "allocating the zoom kernel table")) return rc;
        if (int rc = upload_zoom_screen(p, d, cv, up)) return rc;
    }
    return upload ? up.flush(s) : AMPCOR_HIP_OK;
}

}  // namespace ampcor_hip

extern "C" {

const char *ampcor_hip_last_error(void) { return g_error.c_str(); }

int ampcor_hip_device_count(int *count)
{
    if (!count) return fail(AMPCOR_HIP_EINVAL, "null count");
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        if (e == hipErrorNoDevice) return AMPCOR_HIP_OK;
        return hip_fail(e, "hipGetDeviceCount");
    }
    return AMPCOR_HIP_OK;
}

int ampcor_hip_raster_create(int device, int dtype, const void *host, size_t lines, size_t samples,
                             ampcor_hip_raster **out)
{
    if (!out || !host) return fail(AMPCOR_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (dtype != AMPCOR_HIP_COMPLEX64 && dtype != AMPCOR_HIP_FLOAT32)
        return fail(AMPCOR_HIP_EINVAL, "unknown raster pixel type");
    if (lines == 0 || samples == 0) return fail(AMPCOR_HIP_EINVAL, "empty raster");
    if (int rc = check_device(device)) return rc;
    DeviceGuard guard(device);
    const size_t bytes = lines * samples * (dtype == AMPCOR_HIP_COMPLEX64 ? 8 : 4);
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess)
        return fail(AMPCOR_HIP_ENOMEM, "allocating " + std::to_string(bytes) + " bytes for a raster");
    e = hipMemcpy(d, host, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "uploading a raster");
    }
    auto *r = new (std::nothrow) ampcor_hip_raster{device, dtype, d, lines, samples, true};
    if (!r) {
        (void)hipFree(d);
        return fail(AMPCOR_HIP_ENOMEM, "raster handle");
    }
    *out = r;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_raster_wrap(int device, int dtype, const void *device_ptr, size_t lines,
                           size_t samples, ampcor_hip_raster **out)
{
    if (!out || !device_ptr) return fail(AMPCOR_HIP_EINVAL, "null argument");
    *out = nullptr;
    if (dtype != AMPCOR_HIP_COMPLEX64 && dtype != AMPCOR_HIP_FLOAT32)
        return fail(AMPCOR_HIP_EINVAL, "unknown raster pixel type");
    if (lines == 0 || samples == 0) return fail(AMPCOR_HIP_EINVAL, "empty raster");
    if (int rc = check_device(device)) return rc;
    auto *r = new (std::nothrow) ampcor_hip_raster{device, dtype, device_ptr, lines, samples, false};
    if (!r) return fail(AMPCOR_HIP_ENOMEM, "raster handle");
    *out = r;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_raster_shape(const ampcor_hip_raster *raster, size_t *lines, size_t *samples)
{
    if (!raster || !lines || !samples) return fail(AMPCOR_HIP_EINVAL, "null argument");
    *lines = raster->lines;
    *samples = raster->samples;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_raster_read(const ampcor_hip_raster *raster, size_t offset, size_t count,
                           void *host)
{
    if (!raster || (!host && count)) return fail(AMPCOR_HIP_EINVAL, "null argument");
    const size_t size = raster->lines * raster->samples;
    if (offset > size || count > size - offset)
        return fail(AMPCOR_HIP_EINVAL, "read past the end of the raster");
    if (count == 0) return AMPCOR_HIP_OK;
    const size_t px = raster->dtype == AMPCOR_HIP_COMPLEX64 ? 8 : 4;
    DeviceGuard guard(raster->device);
    HIP_TRY(hipMemcpy(host, static_cast<const char *>(raster->data) + offset * px, count * px,
                      hipMemcpyDeviceToHost),
            "reading raster pixels");
    return AMPCOR_HIP_OK;
}

int ampcor_hip_raster_destroy(ampcor_hip_raster *raster)
{
    if (!raster) return AMPCOR_HIP_OK;
    if (raster->owned && raster->data) {
        DeviceGuard guard(raster->device);
        (void)hipFree(const_cast<void *>(raster->data));
    }
    delete raster;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_plan_create(int device, size_t pairs, size_t ref_lines, size_t ref_samples,
                           size_t tgt_lines, size_t tgt_samples, size_t refine_margin,
                           size_t refine_factor, size_t zoom_factor, int precision,
                           ampcor_hip_plan **out)
{
    if (!out) return fail(AMPCOR_HIP_EINVAL, "null output pointer");
    *out = nullptr;
    auto *p = new (std::nothrow) ampcor_hip_plan();
    if (!p) return fail(AMPCOR_HIP_ENOMEM, "plan");
    if (int rc = plan_configure(p, device, pairs, ref_lines, ref_samples, tgt_lines, tgt_samples,
                                refine_margin, refine_factor, zoom_factor, precision)) {
        delete p;
        return rc;
    }
    DeviceGuard guard(device);
    auto cleanup = [&](int rc) {
        p->release();
        delete p;
        return rc;
    };
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(hip_fail(e, "creating the plan stream"));
    MallocCarver cv;
    if (int rc = plan_build(p, cv, p->stream, true)) return cleanup(rc);
    for (auto &ev : p->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) return cleanup(hip_fail(e, "creating timing events"));
    }
    // the contraction's LDS use can exceed the 64 KiB default
    *out = p;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_plan_destroy(ampcor_hip_plan *plan)
{
    if (!plan) return AMPCOR_HIP_OK;
    {
        DeviceGuard guard(plan->device);
        if (plan->stream) (void)hipStreamSynchronize(plan->stream);
        plan->release();
    }
    delete plan;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_plan_pairs(const ampcor_hip_plan *plan, size_t *pairs)
{
    if (!plan || !pairs) return fail(AMPCOR_HIP_EINVAL, "null argument");
    *pairs = plan->pairs;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_plan_dims(const ampcor_hip_plan *plan, int *coarse_dim, int *refined_dim,
                         int *zoomed_dim)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    if (coarse_dim) *coarse_dim = plan->P;
    if (refined_dim) *refined_dim = plan->R;
    if (zoomed_dim) *zoomed_dim = plan->Z;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_plan_tile_dims(const ampcor_hip_plan *plan, size_t *expanded, size_t *refined_window)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    if (expanded) *expanded = plan->E;
    if (refined_window) *refined_window = plan->Wr;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_add_reference(ampcor_hip_plan *plan, const ampcor_hip_raster *raster, size_t idx,
                             size_t begin_line, size_t begin_sample, size_t end_line,
                             size_t end_sample)
{
    return register_tile(plan, raster, idx, begin_line, begin_sample, end_line, end_sample, true);
}

int ampcor_hip_add_target(ampcor_hip_plan *plan, const ampcor_hip_raster *raster, size_t idx,
                          size_t begin_line, size_t begin_sample, size_t end_line,
                          size_t end_sample)
{
    return register_tile(plan, raster, idx, begin_line, begin_sample, end_line, end_sample, false);
}

int ampcor_hip_add_pairs(ampcor_hip_plan *plan, const ampcor_hip_raster *reference,
                         const ampcor_hip_raster *target, size_t first, size_t count,
                         const int64_t *ref_begin, const int64_t *tgt_begin)
{
    if (!plan || !reference || !target || (count && (!ref_begin || !tgt_begin)))
        return fail(AMPCOR_HIP_EINVAL, "null argument");
    if (first > plan->pairs || count > plan->pairs - first)
        return fail(AMPCOR_HIP_EINVAL, "pair range out of bounds");
    for (size_t k = 0; k < count; ++k) {
        const int64_t rl = ref_begin[2 * k], rs = ref_begin[2 * k + 1];
        const int64_t tl = tgt_begin[2 * k], ts = tgt_begin[2 * k + 1];
        if (rl < 0 || rs < 0 || tl < 0 || ts < 0)
            return fail(AMPCOR_HIP_EINVAL, "negative slice origin at pair " + std::to_string(first + k));
        int rc = register_tile(plan, reference, first + k, rl, rs, rl + plan->C, rs + plan->C, true);
        if (rc) return rc;
        rc = register_tile(plan, target, first + k, tl, ts, tl + plan->W, ts + plan->W, false);
        if (rc) return rc;
    }
    return AMPCOR_HIP_OK;
}

int ampcor_hip_set_timing(ampcor_hip_plan *plan, int enabled)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    plan->timing = enabled != 0;
    return AMPCOR_HIP_OK;
}

}  // extern "C"

namespace ampcor_hip {

int plan_enqueue(ampcor_hip_plan *plan, int N)
{
    hipStream_t s = plan->stream;
    plan->last_n = N;
    if (plan->desc_dirty) {
        HIP_TRY(hipMemcpyAsync(plan->d_desc.ptr, plan->desc.data(), N * sizeof(PairDesc),
                               hipMemcpyHostToDevice, s),
                "uploading pair descriptors");
        plan->desc_dirty = false;
    }
    int stage = 0;
    auto mark = [&](const char *name) -> hipError_t {
        if (!plan->timing || stage >= kStages) return hipSuccess;
        plan->stage_name[stage] = name;
        return hipEventRecord(plan->ev[stage++], s);
    };
#define STEP(call, what)                          \
    do {                                          \
        HIP_TRY(mark(what), "recording an event"); \
        HIP_TRY((call), (what));                  \
    } while (0)

    auto *p = plan;
    const bool fp16 = p->fp16;
    const int qmode = fp16 ? kUpFP16 : kUpBF16;
    if (p->mixed && gather_prep_fits(p->C, p->W)) {
        // the fused normalizer takes the window sums from the quantized target
        STEP(launch_gather_prep_mixed(s, p->d_desc.ptr, N, p->dtype, p->C, p->W,
                                      p->fp8 ? p->fp8_coarse.cs : p->mixed_coarse.cs,
                                      p->fp8 ? kUpFP8 : qmode,
                                      normalize_fuses(p->W, p->C) ? nullptr : p->amp_tgt.ptr,
                                      p->chip_q.ptr, p->tgt_q.ptr, p->mstats.ptr),
             "gather_prep");
    } else {
        STEP(launch_gather_detect(s, p->d_desc.ptr, N, p->dtype, p->C, p->W, p->amp_ref.ptr,
                                  p->amp_tgt.ptr),
             "gather_detect");
        if (p->mixed)
            STEP(launch_prep_mixed(s, p->amp_ref.ptr, p->amp_tgt.ptr, N, p->C, p->W,
                                   p->mixed_coarse.cs, fp16, p->chip_q.ptr, p->tgt_q.ptr,
                                   p->mstats.ptr),
                 "prep");
        else
            STEP(launch_chip_stats(s, p->amp_ref.ptr, N, p->C, p->rT.ptr, p->stats.ptr),
                 "chip_stats");
    }
    // the FP32 coarse stage keeps the separate window sums, whose FP64 order the
    // CPU twin reproduces; the other stages fuse them into the normalizer
    // the packed contractions of small launches in more chip-row ranges (the
    // partial buffers hold the autotuner's largest splits)
    const MixedShape mc = p->mixed ? packed_fill_ranges(p->mixed_coarse, p->C, (int)N,
                                                        kTuneMaxSplitCoarse)
                                   : MixedShape{};
    const MixedShape mr = p->mixed ? packed_fill_ranges(p->mixed_refined, p->Cr, (int)N,
                                                        std::min(kTuneMaxSplitRefined, 4))
                                   : MixedShape{};
    const bool fuse_c = p->mixed && normalize_fuses(p->W, p->C);
    const bool fuse_r = normalize_fuses(p->Wr, p->Cr);
    if (!fuse_c)
        STEP(launch_window_sums(s, p->amp_tgt.ptr, N, p->W, p->C, p->sums.ptr), "window_sums");
    if (p->fp8) {
        const auto *cq8 = reinterpret_cast<const uint8_t *>(p->chip_q.ptr);
        const auto *tq8 = reinterpret_cast<const uint8_t *>(p->tgt_q.ptr);
        STEP(launch_correlate_fp8(s, p->fp8_coarse, cq8, tq8, N, p->C, p->W, p->partial.ptr),
             "correlate");
        STEP(launch_normalize_fused(s, kFuseFP8, tq8, N, p->W, p->C, p->partial.ptr,
                                    p->fp8_coarse.nsplit, nullptr, p->mstats.ptr, p->gamma.ptr),
             "normalize");
    } else if (p->mixed) {
        STEP(launch_correlate_mixed(s, mc, p->chip_q.ptr, p->tgt_q.ptr, N, p->C, p->W, fp16,
                                    p->partial.ptr),
             "correlate");
        if (fuse_c)
            STEP(launch_normalize_fused(s, fp16 ? kFuseFP16 : kFuseBF16, p->tgt_q.ptr, N, p->W,
                                        p->C, p->partial.ptr, mc.nsplit, nullptr, p->mstats.ptr,
                                        p->gamma.ptr),
                 "normalize");
        else
            STEP(launch_normalize_mixed(s, p->partial.ptr, mc.nsplit, p->sums.ptr, p->mstats.ptr, N,
                                        p->C, p->W, p->gamma.ptr),
                 "normalize");
    } else {
        // the exact-order contraction normalizes in its epilogue
        STEP(launch_correlate(s, p->coarse_shape, p->rT.ptr, p->amp_tgt.ptr, N, p->C, p->W,
                              nullptr, p->sums.ptr, p->stats.ptr, p->gamma.ptr),
             "correlate");
    }
    STEP(launch_coarse_peak(s, p->gamma.ptr, N, p->P, p->W, p->E, p->rm, p->rf, p->loc.ptr,
                            p->window.ptr, p->peak.ptr, p->snr.ptr,
                            p->fp8 ? peak_guard_margin(p->C) : 0.0f, p->fp8 ? p->near.ptr : nullptr),
         "coarse_peak");
    // the FP8 surface's near-ties are re-scored in FP64 before the refinement
    // is centred on the peak
    if (p->fp8)
        STEP(launch_peak_guard(s, p->d_desc.ptr, N, p->dtype, p->C, p->W, p->E, p->rm, p->rf,
                               p->gamma.ptr, p->near.ptr, p->loc.ptr, p->window.ptr, p->peak.ptr,
                               p->snr.ptr, p->guard.ptr),
             "peak_guard");
    const bool fused = p->mixed && upsample_fuses_quantization(p->C, p->rf) &&
                       upsample_fuses_quantization(p->E, p->rf);
    // plain mixed forms with the fused upsampling: the refined window sums come
    // from the quantized target, and no FP32 target amplitudes are written
    const bool quant_sums_r = fused && !p->split && normalize_fuses(p->Wr, p->Cr);
    if (fused) {
        STEP(launch_upsample(s, p->d_desc.ptr, N, p->dtype, 0, p->window.ptr, p->C, p->rf,
                             p->U_ref.ptr, p->g_ref.ptr, nullptr, qmode, p->chip_q_r.ptr,
                             p->mixed_refined.cs, p->mstats_r.ptr, p->chip_lo_r.ptr,
                             p->hfrag_ref.ptr, &p->up_scratch),
             "upsample_ref");
        STEP(launch_upsample(s, p->d_desc.ptr, N, p->dtype, 1, p->window.ptr, p->E, p->rf,
                             p->U_tgt.ptr, p->g_tgt.ptr, quant_sums_r ? nullptr : p->amp_tgt_r.ptr,
                             qmode, p->tgt_q_r.ptr, 0, p->mstats_r.ptr, p->tgt_lo_r.ptr,
                             p->hfrag_tgt.ptr, &p->up_scratch),
             "upsample_tgt");
    } else {
        STEP(launch_upsample(s, p->d_desc.ptr, N, p->dtype, 0, p->window.ptr, p->C, p->rf,
                             p->U_ref.ptr, p->g_ref.ptr, p->amp_ref_r.ptr, kUpFP32, nullptr, 0,
                             nullptr, nullptr, nullptr, &p->up_scratch),
             "upsample_ref");
        STEP(launch_upsample(s, p->d_desc.ptr, N, p->dtype, 1, p->window.ptr, p->E, p->rf,
                             p->U_tgt.ptr, p->g_tgt.ptr, p->amp_tgt_r.ptr, kUpFP32, nullptr, 0,
                             nullptr, nullptr, nullptr, &p->up_scratch),
             "upsample_tgt");
        if (p->mixed)
            STEP(launch_prep_mixed(s, p->amp_ref_r.ptr, p->amp_tgt_r.ptr, N, p->Cr, p->Wr,
                                   p->mixed_refined.cs, fp16, p->chip_q_r.ptr, p->tgt_q_r.ptr,
                                   p->mstats_r.ptr, p->chip_lo_r.ptr, p->tgt_lo_r.ptr),
                 "prep_refined");
        else
            STEP(launch_chip_stats(s, p->amp_ref_r.ptr, N, p->Cr, p->rT_r.ptr, p->stats_r.ptr,
                                   false),
                 "chip_stats_refined");
    }
    if (!fuse_r)
        STEP(launch_window_sums(s, p->amp_tgt_r.ptr, N, p->Wr, p->Cr, p->sums_r.ptr),
             "window_sums_refined");
    if (p->mixed) {
        STEP(launch_correlate_mixed(s, mr, p->chip_q_r.ptr, p->tgt_q_r.ptr, N, p->Cr, p->Wr, fp16,
                                    p->partial_r.ptr, p->chip_lo_r.ptr, p->tgt_lo_r.ptr),
             "correlate_refined");
        if (fuse_r)
            STEP(launch_normalize_fused(s, quant_sums_r ? (fp16 ? kFuseFP16 : kFuseBF16) : kFuseMixed,
                                        quant_sums_r ? (const void *)p->tgt_q_r.ptr
                                                     : (const void *)p->amp_tgt_r.ptr,
                                        N, p->Wr, p->Cr, p->partial_r.ptr, mr.nsplit, nullptr,
                                        p->mstats_r.ptr, p->gamma_r.ptr),
                 "normalize_refined");
        else
            STEP(launch_normalize_mixed(s, p->partial_r.ptr, mr.nsplit, p->sums_r.ptr,
                                        p->mstats_r.ptr, N, p->Cr, p->Wr, p->gamma_r.ptr),
                 "normalize_refined");
    } else {
        STEP(launch_correlate(s, p->refined_shape, p->rT_r.ptr, p->amp_tgt_r.ptr, N, p->Cr, p->Wr,
                              p->partial_r.ptr),
             "correlate_refined");
        if (fuse_r)
            STEP(launch_normalize_fused(s, kFuseFP32, p->amp_tgt_r.ptr, N, p->Wr, p->Cr,
                                        p->partial_r.ptr, p->refined_shape.nsplit, p->stats_r.ptr,
                                        nullptr, p->gamma_r.ptr),
                 "normalize_refined");
        else
            STEP(launch_normalize(s, p->partial_r.ptr, p->refined_shape.nsplit, p->sums_r.ptr,
                                  p->stats_r.ptr, N, p->Cr, p->Wr, p->gamma_r.ptr),
                 "normalize_refined");
    }
    STEP(launch_zoom(s, p->gamma_r.ptr, N, p->R, p->zf, p->D.ptr, p->zfrag.ptr, p->zfragp.ptr,
                     p->zoom_bounds, p->window.ptr, p->peak.ptr, p->m, p->rf, p->offsets.ptr,
                     p->peak_r.ptr, p->zoom_loc.ptr, p->zoom_scratch(), p->zstats.ptr),
         "zoom");
#undef STEP
    plan->stages_run = stage;
    if (plan->timing) HIP_TRY(mark("end"), "recording an event");
    return AMPCOR_HIP_OK;
}

}  // namespace ampcor_hip

extern "C" {

int ampcor_hip_adjust(ampcor_hip_plan *plan, float *offsets)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    for (size_t k = 0; k < plan->pairs; ++k)
        if (!plan->have_ref[k] || !plan->have_tgt[k])
            return fail(AMPCOR_HIP_ESTATE, "pair " + std::to_string(k) +
                                               " has no reference or target tile registered");
    DeviceGuard guard(plan->device);
    hipStream_t s = plan->stream;
    const int N = (int)plan->pairs;
    if (int rc = plan_enqueue(plan, N)) return rc;
    const int stages = plan->stages_run;
    const size_t bytes = 2 * (size_t)N * sizeof(float);
    if (offsets) {
        if (!plan->offsets_host)
            HIP_TRY(hipHostMalloc((void **)&plan->offsets_host, 2 * plan->pairs * sizeof(float),
                                  hipHostMallocDefault),
                    "allocating the pinned offsets buffer");
        HIP_TRY(hipMemcpyAsync(plan->offsets_host, plan->offsets.ptr, bytes, hipMemcpyDeviceToHost,
                               s),
                "copying offsets to the host");
    }
    HIP_TRY(hipStreamSynchronize(s), "adjust");
    if (offsets) std::memcpy(offsets, plan->offsets_host, bytes);
    if (plan->timing) {
        for (int k = 0; k < stages; ++k)
            HIP_TRY(hipEventElapsedTime(&plan->stage_ms[k], plan->ev[k], plan->ev[k + 1]),
                    "reading stage times");
    }
    plan->adjusted = true;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_stage_times(ampcor_hip_plan *plan, int capacity, int *count, float *ms,
                           const char **names)
{
    if (!plan || !count) return fail(AMPCOR_HIP_EINVAL, "null argument");
    *count = plan->stages_run;
    if (!plan->timing || !plan->adjusted || plan->stages_run == 0)
        return fail(AMPCOR_HIP_ESTATE, "no timed adjust has run on this plan");
    for (int k = 0; k < plan->stages_run && k < capacity; ++k) {
        if (ms) ms[k] = plan->stage_ms[k];
        if (names) names[k] = plan->stage_name[k];
    }
    return AMPCOR_HIP_OK;
}

static int copy_out(ampcor_hip_plan *plan, void *dst, const void *src, size_t bytes)
{
    if (!dst) return AMPCOR_HIP_OK;
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "copying a product to the host");
    return AMPCOR_HIP_OK;
}

int ampcor_hip_products(ampcor_hip_plan *plan, float *offsets, float *coarse_peak,
                        float *refined_peak, float *snr)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    if (!plan->adjusted) return fail(AMPCOR_HIP_ESTATE, "products requested before adjust");
    DeviceGuard guard(plan->device);
    const size_t N = plan->pairs;
    if (int rc = copy_out(plan, offsets, plan->offsets.ptr, 2 * N * sizeof(float))) return rc;
    if (int rc = copy_out(plan, coarse_peak, plan->peak.ptr, N * sizeof(float))) return rc;
    if (int rc = copy_out(plan, refined_peak, plan->peak_r.ptr, N * sizeof(float))) return rc;
    return copy_out(plan, snr, plan->snr.ptr, N * sizeof(float));
}

int ampcor_hip_stage_gamma(ampcor_hip_plan *plan, float *gamma)
{
    if (!plan || !gamma) return fail(AMPCOR_HIP_EINVAL, "null argument");
    if (!plan->adjusted) return fail(AMPCOR_HIP_ESTATE, "stage requested before adjust");
    DeviceGuard guard(plan->device);
    return copy_out(plan, gamma, plan->gamma.ptr, plan->pairs * plan->P * plan->P * sizeof(float));
}

int ampcor_hip_stage_gamma_refined(ampcor_hip_plan *plan, float *gamma)
{
    if (!plan || !gamma) return fail(AMPCOR_HIP_EINVAL, "null argument");
    if (!plan->adjusted) return fail(AMPCOR_HIP_ESTATE, "stage requested before adjust");
    DeviceGuard guard(plan->device);
    return copy_out(plan, gamma, plan->gamma_r.ptr, plan->pairs * plan->R * plan->R * sizeof(float));
}

int ampcor_hip_stage_upsample(ampcor_hip_plan *plan, int dtype, const void *tiles, size_t count,
                              float *amp)
{
    if (!plan || !tiles || !amp) return fail(AMPCOR_HIP_EINVAL, "null argument");
    if (dtype != AMPCOR_HIP_COMPLEX64 && dtype != AMPCOR_HIP_FLOAT32)
        return fail(AMPCOR_HIP_EINVAL, "unknown raster pixel type");
    if (count == 0 || count > plan->pairs)
        return fail(AMPCOR_HIP_EINVAL, "tile count must be in [1, pairs]");
    DeviceGuard guard(plan->device);
    hipStream_t s = plan->stream;
    const int E = plan->E, M = plan->Wr;
    const size_t px = dtype == AMPCOR_HIP_COMPLEX64 ? 8 : 4, row = count * E * px;
    // the tiles side by side in one raster of E lines, each read as a target
    // window at its own corner with the nudged window at (0, 0)
    std::vector<unsigned char> host((size_t)E * row);
    for (size_t k = 0; k < count; ++k)
        for (int i = 0; i < E; ++i)
            std::memcpy(host.data() + i * row + k * E * px,
                        static_cast<const unsigned char *>(tiles) + (k * E + i) * E * px, E * px);
    void *raster = nullptr;
    HIP_TRY(hipMalloc(&raster, host.size()), "allocating the tile raster");
    std::vector<PairDesc> desc(count);
    for (size_t k = 0; k < count; ++k)
        desc[k] = PairDesc{raster, raster, (int32_t)(count * E), (int32_t)(count * E), 0, 0, 0,
                           (int32_t)(k * E)};
    int rc = AMPCOR_HIP_OK;
    hipError_t e = hipMemcpyAsync(raster, host.data(), host.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(plan->d_desc.ptr, desc.data(), count * sizeof(PairDesc),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(plan->window.ptr, 0, count * sizeof(int2), s);
    // the target upsampling of the plan's adjust: quantized outputs for the plain
    // mixed forms (their FP16 matrix-core path), FP32 amplitudes otherwise
    const bool quant = plan->mixed && upsample_fuses_quantization(plan->E, plan->rf);
    if (e == hipSuccess)
        e = quant ? launch_upsample(s, plan->d_desc.ptr, (int)count, dtype, 1, plan->window.ptr,
                                    plan->E, plan->rf, plan->U_tgt.ptr, plan->g_tgt.ptr,
                                    plan->amp_tgt_r.ptr, plan->fp16 ? kUpFP16 : kUpBF16,
                                    plan->tgt_q_r.ptr, 0, plan->mstats_r.ptr, plan->tgt_lo_r.ptr,
                                    plan->hfrag_tgt.ptr, &plan->up_scratch)
                  : launch_upsample(s, plan->d_desc.ptr, (int)count, dtype, 1, plan->window.ptr,
                                    plan->E, plan->rf, plan->U_tgt.ptr, plan->g_tgt.ptr,
                                    plan->amp_tgt_r.ptr, kUpFP32, nullptr, 0, nullptr, nullptr,
                                    nullptr, &plan->up_scratch);
    if (e == hipSuccess)
        e = hipMemcpyAsync(amp, plan->amp_tgt_r.ptr, count * (size_t)M * M * sizeof(float),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "upsampling tiles");
    (void)hipStreamSynchronize(s);
    (void)hipFree(raster);
    // the plan's own descriptors go up again on the next adjust
    plan->desc_dirty = true;
    plan->adjusted = false;
    return rc;
}

int ampcor_hip_stage_window_sums(ampcor_hip_plan *plan, int stage, const float *windows, int w,
                                 size_t count, double *sums)
{
    if (!plan || !windows || !sums) return fail(AMPCOR_HIP_EINVAL, "null argument");
    if (stage != 0 && stage != 1) return fail(AMPCOR_HIP_EINVAL, "stage must be 0 or 1");
    if (count == 0 || count > plan->pairs)
        return fail(AMPCOR_HIP_EINVAL, "count must be in [1, pairs]");
    const int C = stage ? plan->Cr : plan->C, W = stage ? plan->Wr : plan->W, P = W - C + 1;
    if (w != W) return fail(AMPCOR_HIP_EINVAL, "window size is not the stage's");
    DeviceGuard guard(plan->device);
    hipStream_t s = plan->stream;
    float *amp = stage ? plan->amp_tgt_r.ptr : plan->amp_tgt.ptr;
    double2 *out = stage ? plan->sums_r.ptr : plan->sums.ptr;
    hipError_t e = hipMemcpyAsync(amp, windows, count * (size_t)W * W * sizeof(float),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_window_sums(s, amp, (int)count, W, C, out);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sums, out, count * (size_t)P * P * sizeof(double2),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(s);
        return hip_fail(e, "running the window sums");
    }
    return AMPCOR_HIP_OK;
}

int ampcor_hip_stage_amplitudes(ampcor_hip_plan *plan, int stage, int c, int w, float *chips,
                                float *windows)
{
    if (!plan || !chips || !windows) return fail(AMPCOR_HIP_EINVAL, "null argument");
    if (stage != 0 && stage != 1) return fail(AMPCOR_HIP_EINVAL, "stage must be 0 or 1");
    const int C = stage ? plan->Cr : plan->C, W = stage ? plan->Wr : plan->W;
    if (c != C || w != W) return fail(AMPCOR_HIP_EINVAL, "chip / window sizes are not the stage's");
    if (plan->mixed)
        return fail(AMPCOR_HIP_EINVAL, "the detected amplitudes are kept by FP32 plans only");
    if (plan->last_n <= 0) return fail(AMPCOR_HIP_ESTATE, "no adjust has run");
    DeviceGuard guard(plan->device);
    hipStream_t s = plan->stream;
    const size_t n = (size_t)plan->last_n;
    const float *ca = stage ? plan->amp_ref_r.ptr : plan->amp_ref.ptr;
    const float *wa = stage ? plan->amp_tgt_r.ptr : plan->amp_tgt.ptr;
    hipError_t e = hipMemcpyAsync(chips, ca, n * C * C * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(windows, wa, n * W * W * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "copying the amplitudes");
    return AMPCOR_HIP_OK;
}

int ampcor_hip_stage_contract(ampcor_hip_plan *plan, int stage, int form, const uint16_t *chip,
                              int c, const uint16_t *window, int w, size_t count,
                              float *numerators)
{
    if (!plan || !chip || !window || !numerators) return fail(AMPCOR_HIP_EINVAL, "null argument");
    if (!plan->mixed || plan->split || plan->fp8)
        return fail(AMPCOR_HIP_EINVAL, "stage_contract runs the plain BF16 / FP16 forms");
    if (stage != 0 && stage != 1) return fail(AMPCOR_HIP_EINVAL, "stage must be 0 or 1");
    if (count == 0 || count > plan->pairs)
        return fail(AMPCOR_HIP_EINVAL, "count must be in [1, pairs]");
    const int C = stage ? plan->Cr : plan->C, W = stage ? plan->Wr : plan->W, P = W - C + 1;
    if (c != C || w != W) return fail(AMPCOR_HIP_EINVAL, "chip / window sizes are not the stage's");
    const MixedShape cur = stage ? plan->mixed_refined : plan->mixed_coarse;
    MixedShape sh{};
    switch (form) {
    case 0: sh = cur; break;
    case 1: sh = choose_mixed_shape(C, P); break;
    case 2: sh = choose_mixed_stream_shape(C, P); break;
    case 3: sh = choose_mixed_packed_shape(C, P); break;
    case 4: sh = choose_mixed_packed_shape(C, P, 0, 0, false, 2); break;
    case 5: sh = choose_mixed_packed_shape(C, P, 0, 0, false, 4); break;
    default: return fail(AMPCOR_HIP_EINVAL, "form must be 0 .. 5");
    }
    if (!sh.nk) return fail(AMPCOR_HIP_EINVAL, "the form does not apply to this geometry");
    // the plan's buffers hold the partials of at most its own splits
    const size_t maxsplit = stage ? kTuneMaxSplitRefined : kTuneMaxSplitCoarse;
    if ((size_t)sh.nsplit > std::max<size_t>(maxsplit, cur.nsplit))
        return fail(AMPCOR_HIP_EINVAL, "the form's chip-row split exceeds the plan's buffers");
    DeviceGuard guard(plan->device);
    hipStream_t s = plan->stream;
    uint16_t *cq = stage ? plan->chip_q_r.ptr : plan->chip_q.ptr;
    uint16_t *tq = stage ? plan->tgt_q_r.ptr : plan->tgt_q.ptr;
    float *partial = stage ? plan->partial_r.ptr : plan->partial.ptr;
    // chip rows in the contraction's layout: 16 zeros, the C values, zeros to cs
    std::vector<uint16_t> hc(count * C * (size_t)sh.cs, 0);
    for (size_t k = 0; k < count; ++k)
        for (int i = 0; i < C; ++i)
            std::memcpy(hc.data() + (k * C + i) * sh.cs + 16, chip + (k * C + i) * C,
                        C * sizeof(uint16_t));
    const size_t pp = (size_t)P * P;
    std::vector<float> part(count * sh.nsplit * pp);
    int rc = AMPCOR_HIP_OK;
    hipError_t e = hipMemcpyAsync(cq, hc.data(), hc.size() * sizeof(uint16_t),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(tq, window, count * (size_t)W * W * sizeof(uint16_t),
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_correlate_mixed(s, sh, cq, tq, (int)count, C, W, plan->fp16, partial);
    if (e == hipSuccess)
        e = hipMemcpyAsync(part.data(), partial, part.size() * sizeof(float),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "running the contraction");
    (void)hipStreamSynchronize(s);
    if (rc == AMPCOR_HIP_OK) {
        // chip-row ranges in order, as the normalizers sum them
        for (size_t k = 0; k < count; ++k)
            for (size_t q = 0; q < pp; ++q) {
                float acc = part[(k * sh.nsplit) * pp + q];
                for (int r = 1; r < sh.nsplit; ++r) acc += part[(k * sh.nsplit + r) * pp + q];
                numerators[k * pp + q] = acc;
            }
    }
    plan->adjusted = false;
    return rc;
}

int ampcor_hip_stage_zoom(ampcor_hip_plan *plan, const float *surfaces)
{
    if (!plan || !surfaces) return fail(AMPCOR_HIP_EINVAL, "null argument");
    DeviceGuard guard(plan->device);
    hipStream_t s = plan->stream;
    const size_t N = plan->pairs, RR = (size_t)plan->R * plan->R;
    HIP_TRY(hipMemcpyAsync(plan->gamma_r.ptr, surfaces, N * RR * sizeof(float),
                           hipMemcpyHostToDevice, s),
            "uploading refined surfaces");
    // window corner (0, 0) and a finite coarse peak: offsets = (p + dp) / (rf zf) - m
    HIP_TRY(hipMemsetAsync(plan->window.ptr, 0, N * sizeof(int2), s), "clearing windows");
    HIP_TRY(hipMemsetAsync(plan->peak.ptr, 0, N * sizeof(float), s), "clearing peaks");
    HIP_TRY(launch_zoom(s, plan->gamma_r.ptr, (int)N, plan->R, plan->zf, plan->D.ptr,
                        plan->zfrag.ptr, plan->zfragp.ptr, plan->zoom_bounds, plan->window.ptr,
                        plan->peak.ptr, plan->m, plan->rf, plan->offsets.ptr, plan->peak_r.ptr,
                        plan->zoom_loc.ptr, plan->zoom_scratch(), plan->zstats.ptr),
            "zoom");
    HIP_TRY(hipStreamSynchronize(s), "zoom");
    plan->adjusted = true;
    return AMPCOR_HIP_OK;
}

int ampcor_hip_zoom_stats(ampcor_hip_plan *plan, int64_t *cells, int64_t *pairs,
                          int64_t *searched)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    if (!plan->adjusted) return fail(AMPCOR_HIP_ESTATE, "stage requested before adjust");
    unsigned c[3] = {0, 0, 0};
    DeviceGuard guard(plan->device);
    if (int rc = copy_out(plan, c, plan->zstats.ptr, sizeof(c))) return rc;
    if (cells) *cells = c[0];
    if (pairs) *pairs = c[1];
    if (searched) *searched = c[2];
    return AMPCOR_HIP_OK;
}

int ampcor_hip_fp8_guard_stats(ampcor_hip_plan *plan, int64_t *rescored, int64_t *moved,
                               int64_t *overflow)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    if (!plan->adjusted) return fail(AMPCOR_HIP_ESTATE, "stage requested before adjust");
    unsigned c[3] = {0, 0, 0};
    if (plan->fp8) {
        DeviceGuard guard(plan->device);
        if (int rc = copy_out(plan, c, plan->guard.ptr, sizeof(c))) return rc;
    }
    if (rescored) *rescored = c[0];
    if (moved) *moved = c[1];
    if (overflow) *overflow = c[2];
    return AMPCOR_HIP_OK;
}

int ampcor_hip_stage_locations(ampcor_hip_plan *plan, int32_t *coarse, int32_t *window,
                               int32_t *zoomed)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    if (!plan->adjusted) return fail(AMPCOR_HIP_ESTATE, "stage requested before adjust");
    DeviceGuard guard(plan->device);
    const size_t bytes = plan->pairs * sizeof(int2);
    if (int rc = copy_out(plan, coarse, plan->loc.ptr, bytes)) return rc;
    if (int rc = copy_out(plan, window, plan->window.ptr, bytes)) return rc;
    return copy_out(plan, zoomed, plan->zoom_loc.ptr, bytes);
}

// Times candidate tilings of the two contractions, each followed by its
// normalizer, on the data of the last adjust and keeps the fastest.  The FP32 coarse contraction only varies its
// register block, which leaves its arithmetic order (and results) unchanged.
int ampcor_hip_plan_autotune(ampcor_hip_plan *plan, int reps)
{
    if (!plan) return fail(AMPCOR_HIP_EINVAL, "null plan");
    if (!plan->adjusted) return fail(AMPCOR_HIP_ESTATE, "autotune needs a prior adjust");
    if (reps < 1) reps = 1;
    DeviceGuard guard(plan->device);
    hipStream_t s = plan->stream;
    const int N = plan->last_n > 0 ? plan->last_n : (int)plan->pairs;
    const bool fp16 = plan->fp16;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "creating tuning events");
    HIP_TRY(hipEventCreate(&e1), "creating tuning events");
    auto timeit = [&](auto launch, float *ms) -> int {
        HIP_TRY(launch(), "tuning warm-up");
        HIP_TRY(hipEventRecord(e0, s), "tuning");
        for (int r = 0; r < reps; ++r) HIP_TRY(launch(), "tuning launch");
        HIP_TRY(hipEventRecord(e1, s), "tuning");
        HIP_TRY(hipEventSynchronize(e1), "tuning");
        HIP_TRY(hipEventElapsedTime(ms, e0, e1), "tuning");
        return AMPCOR_HIP_OK;
    };
    // the normalizer reads every partial of the split contraction: candidate
    // splits are timed together with it, launched as adjust launches it
    const bool fuse_c = plan->mixed && normalize_fuses(plan->W, plan->C);
    const bool fuse_r = normalize_fuses(plan->Wr, plan->Cr);
    const bool quant_sums_r = plan->mixed && !plan->split && fuse_r &&
                              upsample_fuses_quantization(plan->C, plan->rf) &&
                              upsample_fuses_quantization(plan->E, plan->rf);
    auto normalize_coarse = [&](int nsplit) -> hipError_t {
        if (plan->fp8)
            return launch_normalize_fused(s, kFuseFP8, plan->tgt_q.ptr, N, plan->W, plan->C,
                                          plan->partial.ptr, nsplit, nullptr, plan->mstats.ptr,
                                          plan->gamma.ptr);
        if (fuse_c)
            return launch_normalize_fused(s, fp16 ? kFuseFP16 : kFuseBF16, plan->tgt_q.ptr, N,
                                          plan->W, plan->C, plan->partial.ptr, nsplit, nullptr,
                                          plan->mstats.ptr, plan->gamma.ptr);
        return launch_normalize_mixed(s, plan->partial.ptr, nsplit, plan->sums.ptr,
                                      plan->mstats.ptr, N, plan->C, plan->W, plan->gamma.ptr);
    };
    auto normalize_refined = [&](int nsplit) -> hipError_t {
        if (!plan->mixed)
            return fuse_r ? launch_normalize_fused(s, kFuseFP32, plan->amp_tgt_r.ptr, N, plan->Wr,
                                                   plan->Cr, plan->partial_r.ptr, nsplit,
                                                   plan->stats_r.ptr, nullptr, plan->gamma_r.ptr)
                          : launch_normalize(s, plan->partial_r.ptr, nsplit, plan->sums_r.ptr,
                                             plan->stats_r.ptr, N, plan->Cr, plan->Wr,
                                             plan->gamma_r.ptr);
        if (fuse_r)
            return launch_normalize_fused(
                s, quant_sums_r ? (fp16 ? kFuseFP16 : kFuseBF16) : kFuseMixed,
                quant_sums_r ? (const void *)plan->tgt_q_r.ptr : (const void *)plan->amp_tgt_r.ptr,
                N, plan->Wr, plan->Cr, plan->partial_r.ptr, nsplit, nullptr, plan->mstats_r.ptr,
                plan->gamma_r.ptr);
        return launch_normalize_mixed(s, plan->partial_r.ptr, nsplit, plan->sums_r.ptr,
                                      plan->mstats_r.ptr, N, plan->Cr, plan->Wr, plan->gamma_r.ptr);
    };
    int rc = AMPCOR_HIP_OK;
    if (!plan->mixed) {
        // Back-to-back launches of one FP32 contraction time it some 10 %
        // slower than the same launch among the other stages, and can rank
        // close candidates differently: a candidate replaces the plan's current
        // tiling, timed first, only when it is more than 5 % faster.
        auto coarse = [&](const CorrelateShape &sh) {
            return launch_correlate(s, sh, plan->rT.ptr, plan->amp_tgt.ptr, N, plan->C, plan->W,
                                    nullptr, plan->sums.ptr, plan->stats.ptr, plan->gamma.ptr);
        };
        CorrelateShape pick = plan->coarse_shape;
        float best = 0.0f;
        rc = timeit([&] { return coarse(pick); }, &best);
        best *= 0.95f;
        for (int bu : {4, 8, 16})
            for (int jb : {0, 16, 8, 4}) {
                if (rc) continue;
                CorrelateShape sh = choose_correlate_shape(plan->C, plan->P, true, bu, 0, jb);
                if (jb && sh.jblock != jb) continue;   // (the default block already)
                float ms;
                if ((rc = timeit([&] { return coarse(sh); }, &ms))) break;
                if (ms < best) best = ms, pick = sh;
            }
        plan->coarse_shape = pick;
        auto refined = [&](const CorrelateShape &sh) {
            hipError_t e = launch_correlate(s, sh, plan->rT_r.ptr, plan->amp_tgt_r.ptr, N,
                                            plan->Cr, plan->Wr, plan->partial_r.ptr);
            return e != hipSuccess ? e : normalize_refined(sh.nsplit);
        };
        pick = plan->refined_shape;
        if (!rc) rc = timeit([&] { return refined(pick); }, &best);
        best *= 0.95f;
        for (int bu : {4, 5, 8, 16})
            for (int ns : {1, 2, 4, 8, 16})
                for (int jb : {0, 16, 8, 4}) {
                    if (rc || plan->Cr % ns || plan->Cr / ns < 8) continue;
                    CorrelateShape sh =
                        choose_correlate_shape(plan->Cr, plan->R, false, bu, ns, jb);
                    if (sh.lds_bytes > 150 * 1024 || (jb && sh.jblock != jb)) continue;
                    float ms;
                    if ((rc = timeit([&] { return refined(sh); }, &ms))) break;
                    if (ms < best) best = ms, pick = sh;
                }
        plan->refined_shape = pick;
    } else {
        struct Stage {
            MixedShape *shape;
            int C, W, P, maxsplit;
            uint16_t *cq, *tq, *cl, *tl;
            float *partial;
            bool split;
        } stages[] = {{&plan->mixed_coarse, plan->C, plan->W, plan->P, kTuneMaxSplitCoarse,
                       plan->chip_q.ptr, plan->tgt_q.ptr, nullptr, nullptr, plan->partial.ptr, false},
                      {&plan->mixed_refined, plan->Cr, plan->Wr, plan->R, kTuneMaxSplitRefined,
                       plan->chip_q_r.ptr, plan->tgt_q_r.ptr, plan->chip_lo_r.ptr, plan->tgt_lo_r.ptr,
                       plan->partial_r.ptr, plan->split}};
        for (auto &st : stages) {
            if (plan->fp8 && st.shape == &plan->mixed_coarse) continue;   // tuned below
            float best = 1e30f;
            const MixedShape orig = *st.shape;
            MixedShape pick = orig;
            bool moved = false;
            const bool coarse = st.shape == &plan->mixed_coarse;
            auto time_shape = [&](const MixedShape &sh, float *ms) {
                return timeit([&] {
                    hipError_t e = launch_correlate_mixed(s, sh, st.cq, st.tq, N, st.C, st.W, fp16,
                                                          st.partial, st.cl, st.tl);
                    if (e != hipSuccess) return e;
                    return coarse ? normalize_coarse(sh.nsplit) : normalize_refined(sh.nsplit);
                }, ms);
            };
            auto consider = [&](const MixedShape &sh) {
                float ms;
                if ((rc = time_shape(sh, &ms))) return;
                if (ms < best) best = ms, pick = sh, moved = true;
            };
            // as for the FP32 form: the current tiling first, and a candidate
            // replaces it only when it is more than 5 % faster
            if ((rc = time_shape(orig, &best))) break;
            best *= 0.95f;
            // v-tiles per workgroup: the widest, and 3 (which divides some tile
            // counts more evenly, e.g. 9 = 3 x 3 rather than 5 + 5)
            for (int vt : {0, 3})
            for (int ns = 1; ns <= st.maxsplit && !rc; ns *= 2) {
                MixedShape sh = choose_mixed_shape(st.C, st.P, ns, st.split, vt);
                if (vt && sh.wg_vt == choose_mixed_shape(st.C, st.P, ns, st.split).wg_vt) continue;
                if (sh.nsplit > st.maxsplit || sh.lds_bytes > 150 * 1024) continue;
                consider(sh);
            }
            // the streaming form (plain operands): all chip rows per workgroup
            for (int vt : {0, 3}) {
                if (st.split || rc) break;
                MixedShape sh = choose_mixed_stream_shape(st.C, st.P, vt);
                if (!sh.nk || (vt && sh.wg_vt == choose_mixed_stream_shape(st.C, st.P).wg_vt))
                    continue;
                consider(sh);
            }
            // the packed form (P = 32 NB + 1) at both register budgets, whole
            // or in chip-row ranges (more workgroups where a pair has few waves)
            for (int wpe : {3, 2})
            for (int ns = 1; ns <= std::min(st.maxsplit, 4) && !rc; ns *= 2) {
                MixedShape sh = choose_mixed_packed_shape(st.C, st.P, 0, wpe, st.split, ns);
                if (sh.nk) consider(sh);
            }
            // a close race timed once can go either way from run to run: the
            // winner and the current tiling are timed again, in turn, and the
            // winner is kept only if it is still more than 5 % faster
            if (moved && !rc) {
                float a = 0.0f, b = 0.0f;
                for (int k = 0; k < 2 && !rc; ++k) {
                    float ms;
                    if ((rc = time_shape(orig, &ms))) break;
                    a += ms;
                    if ((rc = time_shape(pick, &ms))) break;
                    b += ms;
                }
                if (!(b < 0.95f * a)) pick = orig;
            }
            *st.shape = pick;
        }
        if (plan->fp8 && !rc) {
            float best = 1e30f;
            Fp8Shape pick = plan->fp8_coarse;
            const auto *cq8 = reinterpret_cast<const uint8_t *>(plan->chip_q.ptr);
            const auto *tq8 = reinterpret_cast<const uint8_t *>(plan->tgt_q.ptr);
            // the current tiling first; a candidate must be more than 5 % faster
            rc = timeit([&] {
                hipError_t e = launch_correlate_fp8(s, pick, cq8, tq8, N, plan->C, plan->W,
                                                   plan->partial.ptr);
                return e != hipSuccess ? e : normalize_coarse(pick.nsplit);
            }, &best);
            best *= 0.95f;
            for (int vt : {0, 3})
            for (int ns = 1; ns <= kTuneMaxSplitCoarse && !rc; ns *= 2) {
                Fp8Shape sh = choose_fp8_shape(plan->C, plan->P, ns, vt);
                if (vt && sh.wg_vt == choose_fp8_shape(plan->C, plan->P, ns).wg_vt) continue;
                if (sh.nsplit > kTuneMaxSplitCoarse || sh.lds_bytes > 150 * 1024) continue;
                float ms;
                if ((rc = timeit([&] {
                         hipError_t e = launch_correlate_fp8(s, sh, cq8, tq8, N, plan->C, plan->W,
                                                             plan->partial.ptr);
                         return e != hipSuccess ? e : normalize_coarse(sh.nsplit);
                     }, &ms)))
                    break;
                if (ms < best) best = ms, pick = sh;
            }
            plan->fp8_coarse = pick;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    plan->adjusted = false;  // the tuned stages overwrote intermediate buffers
    return rc;
}

}  // extern "C"

namespace ampcor_hip {

size_t plan_bytes_per_pair(int C, int W, int rf, int rm, int precision)
{
    const size_t P = W - C + 1, E = C + 2 * rm, Cr = rf * C, Wr = rf * E, R = Wr - Cr + 1;
    size_t b = sizeof(PairDesc) + 4 * (size_t)(C * C) * 2 + 4 * (size_t)W * W + 16 + 16 * P * P +
               4 * P * P * kTuneMaxSplitCoarse + 4 * P * P + 3 * sizeof(int2) + 4 * 4 + 8 +
               4 * Cr * Cr * 2 + 4 * Wr * Wr + 16 + 16 * R * R + 4 * R * R * kTuneMaxSplitRefined +
               4 * R * R;
    if (precision != AMPCOR_HIP_FP32)
        b += 2 * (size_t)C * (C + 48) + 2 * (size_t)W * W + 2 * Cr * (Cr + 48) + 2 * Wr * Wr +
             2 * sizeof(MixedStats);
    if (precision == AMPCOR_HIP_BF16X3 || precision == AMPCOR_HIP_FP16X3)
        b += 2 * Cr * (Cr + 48) + 2 * Wr * Wr;
    if (precision == AMPCOR_HIP_FP8) b += sizeof(int);  // the peak guard's tie counts
    b += sizeof(double) + 4 * sizeof(int32_t) + sizeof(int);  // the zoom kernels' hand-off
    // banded upsampling scratch for tiles above 96 samples
    size_t nmax = 0;
    if (upsample_needs_scratch(C, rf)) nmax = std::max(nmax, (size_t)C);
    if (upsample_needs_scratch((int)E, rf)) nmax = std::max(nmax, E);
    if (nmax) b += sizeof(float2) * nmax * nmax + 2 * sizeof(double2) * ((nmax + 15) / 16);
    return b;
}

}  // namespace ampcor_hip
