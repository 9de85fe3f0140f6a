"""Pins of the CPU oracle (runs without a GPU).

Each oracle function is checked against something it did not compute itself:
the reference's own known-answer fixtures (SumArea closed forms, planted chip,
the tile-correlate closed form, nudge bounds), brute force written differently
(NumPy), closed forms of the mathematics (band-limited exponentials, Dirichlet
sums, NumPy FFT zero-padding), invariances, and planted mistakes that must be
caught.
"""

import glob
import os

import numpy as np
import pytest

from ampcor_amd import synth
from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

EPS32 = np.finfo(np.float32).eps


# ---- SumArea (tests/libampcor/sumarea.cc:44-69, sumarea-sum.cc:47-148) --------------------

def test_sat_of_ones_is_the_area_table():
    S = po.sat(np.ones((5, 5)))
    i, j = np.meshgrid(np.arange(5), np.arange(5), indexing="ij")
    assert np.array_equal(S[1:, 1:], (i + 1) * (j + 1))
    assert np.all(S[0] == 0) and np.all(S[:, 0] == 0)


def test_sat_window_sums_are_areas():
    S = po.sat(np.ones((5, 6)))
    assert po.sat_sum(S, 0, 0, 5, 6) == 30          # whole grid
    assert po.sat_sum(S, 1, 1, 4, 5) == 12          # interior
    assert po.sat_sum(S, 1, 0, 5, 6) == 24          # minus a row
    assert po.sat_sum(S, 0, 1, 5, 6) == 25          # minus a column


def test_sat_matches_brute_force():
    rng = np.random.default_rng(0)
    t = rng.random((17, 23))
    S = po.sat(t)
    for r0, c0, r1, c1 in [(0, 0, 17, 23), (3, 4, 9, 20), (16, 22, 17, 23), (5, 5, 5, 9)]:
        assert po.sat_sum(S, r0, c0, r1, c1) == pytest.approx(t[r0:r1, c0:c1].sum(), rel=1e-13, abs=1e-13)


# ---- NCC ------------------------------------------------------------------------------------

def ncc_numpy(ref, tgt):
    """Brute-force NCC written as explicit window extraction (no SAT)."""
    C, W = ref.shape[0], tgt.shape[0]
    r = ref - ref.mean()
    out = np.empty((W - C + 1, W - C + 1))
    for u in range(W - C + 1):
        for v in range(W - C + 1):
            win = tgt[u:u + C, v:v + C]
            d = win - win.mean()
            out[u, v] = (r * d).sum() / np.sqrt((r * r).sum() * (d * d).sum())
    return out


@pytest.mark.parametrize("C,W", [(2, 4), (4, 8), (8, 14)])
def test_ncc_matches_brute_force(C, W):
    rng = np.random.default_rng(C)
    ref, tgt = rng.random((C, C)), rng.random((W, W))
    np.testing.assert_allclose(po.ncc64(ref, tgt), ncc_numpy(ref, tgt), rtol=1e-12, atol=1e-13)


def test_ncc_is_scale_and_offset_invariant():
    rng = np.random.default_rng(1)
    ref, tgt = rng.random((8, 8)), rng.random((16, 16))
    a = po.ncc64(ref, tgt)
    b = po.ncc64(3.0 * ref + 7.0, 0.5 * tgt - 2.0)
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-12)


def test_ncc_planted_chip_is_one():
    """tests/libampcor_cuda/correlate.cc:283-306: gamma = 1 at the planted placement."""
    rng = np.random.default_rng(2)
    C, m = 16, 4
    chip = np.abs(rng.standard_normal((C, C)))
    for i, j in [(0, 0), (4, 4), (8, 3), (2, 8)]:
        tgt = np.zeros((C + 2 * m, C + 2 * m))
        tgt[i:i + C, j:j + C] = chip
        g = po.ncc64(chip, tgt)
        assert abs(g[i, j] - 1.0) <= 10 * EPS32
        assert po.argmax(g)[:2] == (i, j)


def test_cross_sum_tile_correlate_closed_form():
    """tests/cuda/tile-correlate.cu:359-388: a chip of ones against ones padded by
    zeros gives (C - |m - r|)(C - |m - c|)."""
    C, m = 8, 3
    tgt = np.zeros((C + 2 * m, C + 2 * m))
    tgt[m:m + C, m:m + C] = 1.0
    got = po.cross_sum(np.ones((C, C)), tgt)
    r, c = np.meshgrid(np.arange(2 * m + 1), np.arange(2 * m + 1), indexing="ij")
    assert np.array_equal(got, (C - np.abs(m - r)) * (C - np.abs(m - c)))


def test_ncc_zero_variance_window_is_nan():
    g = po.ncc64(np.arange(16.0).reshape(4, 4), np.ones((6, 6)))
    assert np.all(np.isnan(g))


# ---- argmax and nudge (maxcor.cu:110-129, nudge.cu:113-139) ---------------------------------

def test_argmax_first_max_and_nan_policy():
    g = np.array([[0.1, 0.9, 0.2], [0.9, 0.3, np.nan]], dtype=np.float32)
    assert po.argmax(g)[:2] == (0, 1)                       # first of the tied maxima
    g[0, 0] = np.nan
    assert po.argmax(g)[:2] == (0, 1)                       # NaN at cell 0 never wins
    r, c, pk = po.argmax(np.full((3, 3), np.nan, dtype=np.float32))
    assert (r, c) == (0, 0) and np.isnan(pk)
    g64 = np.zeros((4, 5))
    g64[3, 4] = 2.0
    assert po.argmax(g64) == (3, 4, 2.0)


def test_nudge_closed_forms():
    W, rm, C = 128, 8, 64
    E = C + 2 * rm
    for row in range(0, W - C + 1):
        top, left = po.nudge(row, W - C - row, rm, W, E)
        assert 0 <= top <= W - E and 0 <= left <= W - E
        if rm <= row <= W - E + rm:
            assert top == row - rm                          # interior: identity shift


# ---- upsampling ------------------------------------------------------------------------------

def fft_upsample(x, rf):
    """Zero-padded spectrum with the Nyquist bin split (NumPy FFT)."""
    N = x.shape[0]
    M = rf * N
    X = np.fft.fft2(x)
    Y = np.zeros((M, M), dtype=complex)
    half = N // 2
    idx = list(range(-(N - 1) // 2 if N % 2 else -half + 1, (N - 1) // 2 + 1 if N % 2 else half))
    for a in idx:
        for b in idx:
            Y[a % M, b % M] = X[a % N, b % N]
    if N % 2 == 0:
        # split the Nyquist row/column in half between +N/2 and -N/2
        for b in idx + [half]:
            wb = 0.5 if b == half else 1.0
            for sa in (half, -half):
                Y[sa % M, b % M] += 0.5 * wb * X[half, b % N]
                if b == half:
                    Y[sa % M, -half % M] += 0.25 * X[half, half]
        for a in idx:
            for sb in (half, -half):
                Y[a % M, sb % M] += 0.5 * X[a % N, half]
    return np.fft.ifft2(Y) * rf * rf


@pytest.mark.parametrize("N,rf", [(8, 2), (9, 2), (6, 3), (16, 2)])
def test_upsample_matches_fft_zero_padding(N, rf):
    rng = np.random.default_rng(N + rf)
    x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    np.testing.assert_allclose(po.upsample(x, rf), fft_upsample(x, rf), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("N,rf", [(8, 2), (7, 3), (80, 2)])
def test_upsample_exact_on_band_limited_exponentials(N, rf):
    k, l = (N // 2 - 1, -(N // 2 - 2))
    n = np.arange(N)
    x = np.exp(2j * np.pi * (k * n[:, None] + l * n[None, :]) / N)
    p = np.arange(rf * N) / rf
    want = np.exp(2j * np.pi * (k * p[:, None] + l * p[None, :]) / N)
    np.testing.assert_allclose(po.upsample(x, rf), want, rtol=0, atol=1e-11)


def test_upsample_is_identity_on_the_coarse_subgrid():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))
    assert np.array_equal(po.upsample(x, 2)[::2, ::2], x)


# ---- zoom --------------------------------------------------------------------------------------

def fft_zoom(g, zf):
    R = g.shape[0]
    Z = zf * R
    G = np.fft.fft2(g)
    Y = np.zeros((Z, Z), dtype=complex)
    h = (R - 1) // 2
    for a in range(-h, h + 1):
        for b in range(-h, h + 1):
            Y[a % Z, b % Z] = G[a % R, b % R]
    return np.real(np.fft.ifft2(Y)) * zf * zf


@pytest.mark.parametrize("R,zf", [(9, 4), (33, 8), (5, 3)])
def test_zoom_matches_fft_zoom(R, zf):
    g = np.random.default_rng(R).standard_normal((R, R))
    np.testing.assert_allclose(po.zoom_full(g, zf), fft_zoom(g, zf), rtol=1e-10, atol=1e-11)


def test_zoom_identity_on_subgrid_and_bump_peak():
    R, zf = 33, 8
    g = np.random.default_rng(5).standard_normal((R, R))
    assert np.array_equal(po.zoom_full(g, zf)[::zf, ::zf], g)
    # a band-limited cosine bump peaking at a known off-grid point
    x0, y0 = 14.375, 18.625
    a = np.arange(R)
    f = lambda x, c: np.cos(2 * np.pi * (x - c) / R)
    bump = f(a[:, None], x0) + f(a[None, :], y0)
    z = po.zoom_full(bump, zf)
    p, q = np.unravel_index(np.argmax(z), z.shape)
    assert (p / zf, q / zf) == (x0, y0)


def test_dirichlet_closed_form_against_sum():
    R, zf = 9, 4
    for d in range(-R * zf, R * zf + 1):
        x = d / zf
        k = np.arange(-(R - 1) // 2, (R - 1) // 2 + 1)
        want = np.real(np.exp(2j * np.pi * k * x / R).sum()) / R
        assert po.dirichlet(R, zf, d) == pytest.approx(want, abs=1e-13)


# ---- SNR ---------------------------------------------------------------------------------------

def test_snr_closed_form():
    g = np.full((9, 9), 0.1)
    g[4, 4] = 0.8
    g[3:6, 3:6] = 0.5
    g[4, 4] = 0.8
    # box of radius 1 around the peak excluded: 72 cells of 0.1
    assert po.snr(g, 4, 4, 1) == pytest.approx(0.64 / 0.01, rel=1e-12)
    g[0, 0] = np.nan
    assert po.snr(g, 4, 4, 1) == pytest.approx(0.64 / 0.01, rel=1e-12)


# ---- the whole pipeline --------------------------------------------------------------------------

def test_pipeline_recovers_fourier_shift():
    s = synth.shifted_scene()
    out = po.adjust(s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                    s.refine_factor, s.refine_margin, s.zoom_factor)
    assert s.pairs == 64
    err = np.abs(out["offsets"] - s.truth)
    assert err.max() < 0.05 and np.median(err) < 0.01


def test_twin_agrees_with_fp64_on_coarse_stage():
    s = synth.shifted_scene(seed=9, size=256, grid=(4, 4))
    a = po.adjust(s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor, form="fp32", surfaces=True)
    b = po.adjust(s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor, form="fp64", surfaces=True)
    assert np.array_equal(a["coarse_loc"], b["coarse_loc"])
    np.testing.assert_allclose(a["gamma"], b["gamma"], atol=4e-6)


def test_pipeline_planted_chip_recovers_placement():
    ref, tgt, rb, tb = synth.planted(16, 4, seed=3)
    out = po.adjust(ref, tgt, rb, tb, 16, 24, 2, 4, 4)
    truth = np.array([(i - 4, j - 4) for i in range(9) for j in range(9)], np.float32)
    assert np.array_equal(out["coarse_loc"], truth.astype(np.int32) + 4)
    assert np.max(np.abs(out["offsets"] - truth)) < 0.1


def test_planted_mistake_is_caught():
    """A sign error in the offset assembly (tgt - ref reversed) must fail the shift check."""
    s = synth.shifted_scene(seed=1)
    out = po.adjust(s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                    s.refine_factor, s.refine_margin, s.zoom_factor)
    assert np.max(np.abs(-out["offsets"] - s.truth)) > 1.0


def test_invalid_arguments_are_rejected():
    s = synth.shifted_scene(seed=1, size=128, grid=(2, 2))
    with pytest.raises(ValueError):
        po.adjust(s.reference, s.target, s.ref_begin, s.tgt_begin, 32, 64, 2, 20, 4)  # rm > m
    bad = s.tgt_begin.copy()
    bad[0, 0] = 100
    with pytest.raises(ValueError):
        po.adjust(s.reference, s.target, s.ref_begin, bad, 32, 64, 2, 8, 4)          # off raster


def test_parabola_vertex_of_sampled_quadratics():
    # f(x) = c - a (x - x0)^2 sampled at -1, 0, 1: the vertex is x0 exactly
    for a in (0.3, 2.0):
        for c in (0.0, 0.9):
            for x0 in np.linspace(-0.49, 0.49, 99):
                f = [c - a * (k - x0) ** 2 for k in (-1.0, 0.0, 1.0)]
                assert abs(po.parabola(*f) - x0) < 1e-12
    assert po.parabola(1.0, 1.0, 1.0) == 0.0            # flat: no step
    assert po.parabola(1.0, 0.0, 1.0) == 0.0            # a minimum: no step
    assert po.parabola(0.0, 1.0, 1.5) == 0.5            # vertex at 1.5: clamped
    assert po.parabola(1.5, 1.0, 0.0) == -0.5


def zoom_full_answer(g, zf):
    z = po.zoom_full(np.nan_to_num(g), zf)
    p, q = np.unravel_index(np.argmax(z), z.shape)
    Z = z.shape[0]
    dp = po.parabola(z[p - 1, q], z[p, q], z[p + 1, q]) if 0 < p < Z - 1 else 0.0
    dq = po.parabola(z[p, q - 1], z[p, q], z[p, q + 1]) if 0 < q < Z - 1 else 0.0
    return (int(p), int(q)), z[p, q], (dp, dq)


def test_zoom_search_is_the_full_grid_maximum():
    """The pipeline's zoom search (oracle_zoom_search) returns the first maximum of
    the whole zoomed grid and the parabolas through its neighbours, on the
    golden surfaces, random surfaces and refined surfaces of a known shift."""
    surfaces = []
    for path in sorted(glob.glob(os.path.join(GOLDEN, "zoom_*.npz"))):
        f = np.load(path)
        surfaces.append((f["g"], int(f["zf"])))
    rng = np.random.default_rng(17)
    for R, zf in ((9, 4), (11, 3), (17, 2)):
        for _ in range(4):
            surfaces.append((rng.uniform(-1, 1, (R, R)), zf))
    f = np.load(os.path.join(GOLDEN, "shift_offgrid.npz"))
    C, W, rf, rm, zf = (int(v) for v in f["geometry"])
    out = po.adjust(f["reference"], f["target"], f["ref_begin"][:4], f["tgt_begin"][:4], C, W, rf,
                    rm, zf, surfaces=True)
    surfaces += [(g.astype(np.float64), zf) for g in out["gamma_refined"]]
    far = 0
    for g, zf in surfaces:
        (p, q), pk, (dp, dq) = po.zoom_search(g, zf)
        (ep, eq), epk, (edp, edq) = zoom_full_answer(g, zf)
        assert (p, q) == (ep, eq)
        assert abs(pk - epk) <= 1e-12 * max(1.0, abs(epk))
        assert abs(dp - edp) < 1e-9 and abs(dq - edq) < 1e-9
        a, b = np.unravel_index(np.argmax(g), g.shape)
        far += max(abs(p - zf * a), abs(q - zf * b)) > zf
    # on several of them the maximum is more than a refined sample away from
    # the largest sample: a search near that sample would miss it
    assert far >= 3
