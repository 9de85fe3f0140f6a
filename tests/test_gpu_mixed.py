"""The mixed-precision forms (BF16 / FP16 contraction on MFMA, FP32 accumulate,
FP64 window statistics) against the FP64 oracle.

Bar: refined offsets within 0.01 px of the FP64 oracle; coarse peaks on the
FP64 oracle's placement; correlation peaks within 1e-2 relative.
"""

import numpy as np
import pytest

from ampcor_amd import synth
from oracle import pyoracle
from tests.gpu_helpers import run_hip

pytestmark = pytest.mark.gpu

MIXED_OFFSET_TOL = 0.01  # px vs the FP64 oracle
PEAK_REL_TOL = 1e-2


def fp64(s):
    return pyoracle.adjust(s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                           s.refine_factor, s.refine_margin, s.zoom_factor, form="fp64")


def run(lib, s, precision):
    return run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                   s.refine_factor, s.refine_margin, s.zoom_factor, precision=precision)


@pytest.mark.parametrize("precision", ["bf16", "fp16", "bf16x3", "fp16x3"])
def test_mixed_shifted_scene(lib, precision):
    s = synth.shifted_scene()
    hip, exact = run(lib, s, precision), fp64(s)
    err = np.abs(hip["offsets"] - exact["offsets"])
    print(f"{precision}: max |offset - fp64| = {err.max():.2e} px", flush=True)
    assert err.max() <= MIXED_OFFSET_TOL
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])
    np.testing.assert_allclose(hip["coarse_peak"], exact["coarse_peak"], rtol=PEAK_REL_TOL)
    np.testing.assert_allclose(hip["refined_peak"], exact["refined_peak"], rtol=PEAK_REL_TOL)


@pytest.mark.parametrize("precision", ["bf16", "fp16", "bf16x3", "fp16x3"])
def test_mixed_field_scene(lib, precision):
    s = synth.field_scene(seed=6, lines=1024, samples=1024, grid=(12, 12), range_ramp=3.0,
                          azimuth_quadratic=1.5)
    hip, exact = run(lib, s, precision), fp64(s)
    err = np.abs(hip["offsets"] - exact["offsets"])
    print(f"{precision}: {s.pairs} pairs, max |offset - fp64| = {err.max():.2e} px", flush=True)
    assert err.max() <= MIXED_OFFSET_TOL
    agree = np.mean(np.all(hip["coarse_loc"] == exact["coarse_loc"], axis=1))
    assert agree >= 0.98


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_mixed_planted_chip(lib, precision):
    ref, tgt, rb, tb = synth.planted(32, 8, seed=4)
    hip = run_hip(lib, ref, tgt, rb, tb, 32, 48, 2, 4, 4, precision=precision)
    P = 17
    for i in range(P):
        for j in range(P):
            k = i * P + j
            assert tuple(hip["coarse_loc"][k]) == (i, j)
            assert abs(hip["gamma"][k, i, j] - 1.0) < 1e-2


def test_mixed_float32_rasters(lib):
    rng = np.random.default_rng(12)
    ref = synth.lognormal_texture((384, 384), 0.5, rng)
    tgt = np.real(synth.fourier_shift(ref, -2.5, 1.75)).astype(np.float32)
    tgt = tgt * (1 + 0.05 * rng.standard_normal(tgt.shape)).astype(np.float32)
    s = synth.Scene("amp", ref, tgt, 32, 16, 2, 8, 8, (4, 4))
    s.plan()
    hip, exact = run(lib, s, "fp16"), fp64(s)
    assert np.abs(hip["offsets"] - exact["offsets"]).max() <= MIXED_OFFSET_TOL


def test_split_refinement_is_closer_to_fp64(lib):
    """The hi/lo products of the split form win back accuracy in the refined stage."""
    s = synth.field_scene(seed=7, lines=1024, samples=1024, grid=(12, 12), range_ramp=3.0,
                          azimuth_quadratic=1.5)
    exact = fp64(s)
    plain = np.abs(run(lib, s, "bf16")["offsets"] - exact["offsets"]).max()
    split = np.abs(run(lib, s, "bf16x3")["offsets"] - exact["offsets"]).max()
    print(f"bf16 {plain:.2e} px, bf16x3 {split:.2e} px", flush=True)
    assert split <= plain


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
def test_mixed_window_not_a_multiple_of_8(lib, precision):
    """W = 52: window rows are 4-byte but not 16-byte aligned (the contractions'
    dword staging, the gathers' odd-chunk paths)."""
    s = synth.shifted_scene(chip=32, margin=10, refine_margin=8)
    hip, exact = run(lib, s, precision), fp64(s)
    err = np.abs(hip["offsets"] - exact["offsets"])
    print(f"{precision} W=52: max |offset - fp64| = {err.max():.2e} px", flush=True)
    assert err.max() <= MIXED_OFFSET_TOL
    assert np.mean(np.all(hip["coarse_loc"] == exact["coarse_loc"], axis=1)) >= 0.98


@pytest.fixture(scope="module")
def wide_complex():
    # complex tiles above 96 samples (chip 128, refined window 144): the banded
    # two-pass upsampler in its quantized modes, as the serving stream's chip-128 jobs use it
    s = synth.shifted_scene(seed=17, size=640, grid=(2, 2), chip=128, margin=32,
                            refine_margin=8, zoom_factor=4, shift=(1.3, -2.6))
    return s, fp64(s)


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16", "bf16x3"])
def test_complex_chip_128_banded_upsampler(lib, wide_complex, precision):
    s, exact = wide_complex
    print(f"wide complex tiles: {s.pairs} pairs, chip {s.chip}, window {s.window}, {precision}",
          flush=True)
    hip = run(lib, s, precision)
    err = np.abs(hip["offsets"] - exact["offsets"])
    assert err.max() <= (1e-3 if precision in ("fp32", "bf16x3") else MIXED_OFFSET_TOL)
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])
    assert np.max(np.abs(hip["offsets"] - s.truth)) < 0.05


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_surface_above_144_placements(lib, precision):
    # 145 x 145 placements (margin 72): past the fused normalizer's reach, the
    # mixed coarse stage sums the window in FP64 separately (k_window_sums +
    # k_normalize_mixed), and the coarse peak reads the surface twice
    s = synth.shifted_scene(seed=31, size=512, grid=(2, 2), chip=64, margin=72, refine_margin=8,
                            zoom_factor=4, shift=(5.7, -4.4))
    print(f"145 x 145 placements: {s.pairs} pairs, {precision}", flush=True)
    exact = fp64(s)
    hip = run(lib, s, precision)
    assert np.max(np.abs(hip["offsets"] - exact["offsets"])) <= MIXED_OFFSET_TOL
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])
    np.testing.assert_allclose(hip["coarse_peak"], exact["coarse_peak"], rtol=PEAK_REL_TOL)
