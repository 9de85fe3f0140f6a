"""Parity of the HIP correlator with the CPU oracle (runs on the GPU).

Bar: coarse surfaces, coarse peaks and nudged windows bit-identical to the
oracle's FP32 twin; refined offsets within 1e-3 px, peaks and SNR within 1e-5
relative of the FP64 oracle.
"""

import numpy as np
import pytest

from ampcor_amd import synth
from oracle import pyoracle
from tests.gpu_helpers import run_hip

pytestmark = pytest.mark.gpu

OFFSET_TOL = 1e-3   # px, refined offsets vs the FP64 oracle
REL_TOL = 1e-5      # correlation peaks and SNR vs the FP64 oracle


def oracles(s, reference=None, target=None):
    reference = s.reference if reference is None else reference
    target = s.target if target is None else target
    args = (reference, target, s.ref_begin, s.tgt_begin, s.chip, s.window, s.refine_factor,
            s.refine_margin, s.zoom_factor)
    return (pyoracle.adjust(*args, form="fp32", surfaces=True),
            pyoracle.adjust(*args, form="fp64", surfaces=True))


def check_against_oracles(hip, twin, exact):
    # coarse stage: bit-identical to the FP32 twin
    assert np.array_equal(hip["gamma"].view(np.uint32), twin["gamma"].view(np.uint32))
    assert np.array_equal(hip["coarse_loc"], twin["coarse_loc"])
    assert np.array_equal(hip["window_loc"], twin["window_loc"])
    assert np.array_equal(hip["coarse_peak"].view(np.uint32), twin["coarse_peak"].view(np.uint32))
    # refinement and products: FP64 oracle
    np.testing.assert_allclose(hip["offsets"], exact["offsets"], rtol=0, atol=OFFSET_TOL)
    np.testing.assert_allclose(hip["coarse_peak"], exact["coarse_peak"], rtol=REL_TOL)
    np.testing.assert_allclose(hip["refined_peak"], exact["refined_peak"], rtol=REL_TOL)
    np.testing.assert_allclose(hip["snr"], exact["snr"], rtol=REL_TOL)
    # the refined surfaces themselves, within 1e-5 of each surface's peak
    gr, er = hip["gamma_refined"].astype(np.float64), exact["gamma_refined"].astype(np.float64)
    scale = np.nanmax(np.abs(er), axis=(1, 2))[:, None, None]
    assert np.nanmax(np.abs(gr - er) / scale) <= REL_TOL
    # the tuple adjust returns is the offsets product
    assert np.array_equal(hip["tuple_offsets"], hip["offsets"])


def test_shifted_speckle_matches_oracles(lib):
    s = synth.shifted_scene()
    print(f"shifted scene: {s.pairs} pairs", flush=True)
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)
    # and the known shift is recovered
    assert np.max(np.abs(hip["offsets"] - s.truth)) < 0.05


def test_tuned_plan_matches_oracles(lib):
    """After the autotuner has picked its tilings (every FP32 coarse tiling keeps
    the exact order; the refined one may split the chip columns into 1 .. 16
    ranges), the same bars hold."""
    s = synth.shifted_scene(seed=5)
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor, tune=True)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)


def test_blocks_that_do_not_divide_the_chip(lib, monkeypatch):
    """Five placement rows per lane (the chips of 32 and 64 samples end each
    column's chain with a part of a block): the coarse surfaces stay bitwise
    the FP32 twin's, the refined ones within the FP64 bars."""
    monkeypatch.setenv("AMPCOR_HIP_TILE_COARSE", "5")
    monkeypatch.setenv("AMPCOR_HIP_TILE_REFINED", "5,4")
    s = synth.shifted_scene(seed=6)
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)


@pytest.mark.parametrize("jblock", [4, 7])
def test_chip_columns_in_several_lds_passes(lib, monkeypatch, jblock):
    """The FP32 contractions with the chip columns taken a few at a time (the
    LDS block a ring of target columns; 7 leaves a shorter last pass): coarse
    surfaces bitwise the twin's, refined ones within the FP64 bars."""
    monkeypatch.setenv("AMPCOR_HIP_TILE_COARSE", f"8,{jblock}")
    monkeypatch.setenv("AMPCOR_HIP_TILE_REFINED", f"4,2,{jblock}")
    s = synth.shifted_scene(seed=8)
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)


def test_golden_shift_end_to_end(lib):
    """The committed fixture: every pair recovers the planted Fourier shift."""
    import os
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shift_e2e.npz"))
    C, W, rf, rm, zf = (int(v) for v in f["geometry"])
    hip = run_hip(lib, f["reference"], f["target"], f["ref_begin"], f["tgt_begin"], C, W, rf, rm, zf)
    assert np.max(np.abs(hip["offsets"] - f["shift"])) < 0.05
    exact = pyoracle.adjust(f["reference"], f["target"], f["ref_begin"], f["tgt_begin"], C, W, rf, rm,
                            zf, form="fp64")
    np.testing.assert_allclose(hip["offsets"], exact["offsets"], rtol=0, atol=OFFSET_TOL)


@pytest.mark.parametrize("precision,tol", [("fp32", OFFSET_TOL), ("bf16", 1e-2)])
def test_golden_offgrid_shift(lib, precision, tol):
    """The committed off-grid fixture (+3.31, -1.69 px, not a multiple of the
    1/8 px zoomed cell): the sub-pixel step carries ~0.06 px of each offset."""
    import os
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "shift_offgrid.npz"))
    C, W, rf, rm, zf = (int(v) for v in f["geometry"])
    hip = run_hip(lib, f["reference"], f["target"], f["ref_begin"], f["tgt_begin"], C, W, rf, rm,
                  zf, precision=precision)
    exact = pyoracle.adjust(f["reference"], f["target"], f["ref_begin"], f["tgt_begin"], C, W, rf,
                            rm, zf, form="fp64")
    np.testing.assert_allclose(hip["offsets"], exact["offsets"], rtol=0, atol=tol)
    assert np.max(np.abs(hip["offsets"] - f["shift"])) < 0.03
    assert np.array_equal(hip["zoom_loc"], exact["zoom_loc"])


def test_per_pair_calls_equal_batched(lib):
    s = synth.shifted_scene(seed=3, size=256, grid=(4, 4))
    a = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                s.refine_factor, s.refine_margin, s.zoom_factor, per_pair=True)
    b = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                s.refine_factor, s.refine_margin, s.zoom_factor)
    for k in ("offsets", "gamma", "gamma_refined", "snr"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


@pytest.mark.parametrize("chip,margin", [(16, 4), (32, 8)])
def test_planted_chip(lib, chip, margin):
    """tests/libampcor_cuda/correlate.cc:283-306 and maxcor.cc:296-315: gamma = 1
    at the planted placement (10 eps) and the argmax is exactly that placement."""
    ref, tgt, rb, tb = synth.planted(chip, margin, seed=chip)
    P = 2 * margin + 1
    rm = min(margin, 4)
    hip = run_hip(lib, ref, tgt, rb, tb, chip, chip + 2 * margin, 2, rm, 4)
    eps = np.finfo(np.float32).eps
    for i in range(P):
        for j in range(P):
            k = i * P + j
            assert abs(hip["gamma"][k, i, j] - 1.0) <= 10 * eps
            assert tuple(hip["coarse_loc"][k]) == (i, j)
    # nudge bounds (nudge.cc:201-227)
    E = chip + 2 * rm
    assert np.all(hip["window_loc"] >= 0) and np.all(hip["window_loc"] <= chip + 2 * margin - E)
    # refined offsets agree with the FP64 oracle and land on the planted placement
    exact = pyoracle.adjust(ref, tgt, rb, tb, chip, chip + 2 * margin, 2, rm, 4, form="fp64")
    np.testing.assert_allclose(hip["offsets"], exact["offsets"], rtol=0, atol=OFFSET_TOL)
    truth = np.array([(i - margin, j - margin) for i in range(P) for j in range(P)], np.float32)
    assert np.max(np.abs(hip["offsets"] - truth)) < 0.1


def test_field_scene_subset(lib):
    """Config-2 geometry (chip 64, window 128, zoom 8) on a smaller raster."""
    s = synth.field_scene(seed=5, lines=1024, samples=1024, grid=(12, 12), range_ramp=3.0,
                          azimuth_quadratic=1.5)
    print(f"field subset: {s.pairs} pairs", flush=True)
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)
    assert np.max(np.abs(hip["offsets"] - s.truth)) < 0.1


def test_float32_amplitude_rasters(lib):
    rng = np.random.default_rng(11)
    ref = synth.lognormal_texture((384, 384), 0.5, rng)
    tgt = np.real(synth.fourier_shift(ref, 1.5, 2.25)).astype(np.float32)
    tgt = tgt * (1 + 0.05 * rng.standard_normal(tgt.shape)).astype(np.float32)
    s = synth.Scene("amp", ref, tgt, 32, 16, 2, 8, 8, (4, 4))
    s.plan()
    hip = run_hip(lib, ref, tgt, s.ref_begin, s.tgt_begin, s.chip, s.window, 2, 8, 8)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)


def test_wide_refined_surface_unstaged_zoom(lib):
    """R = 2 rf rm + 1 = 125 with zoom 4: the zoom's Dirichlet weights no longer
    fit its staging budget and are read from global memory (the oracle takes
    R <= 128)."""
    s = synth.shifted_scene(size=384, grid=(3, 3), chip=32, margin=32, refine_factor=2,
                            refine_margin=31, zoom_factor=4)
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)


def test_slc_getitem_and_refine(lib):
    # slc_getitem reads one pixel at a row-major offset (ext/ampcor/slc.cc:105-140);
    # refine is the CPU module's separate call, a no-op after adjust
    rng = np.random.default_rng(12)
    z = (rng.standard_normal((7, 9)) + 1j * rng.standard_normal((7, 9))).astype(np.complex64)
    a = rng.standard_normal((5, 6)).astype(np.float32)
    rz = lib.raster(z, 7, 9)
    ra = lib.raster(a, 5, 6, dtype="float32")
    for off in (0, 1, 8, 9, 40, 62):
        assert lib.slc_getitem(rz, off) == complex(z.reshape(-1)[off])
    for off in (0, 7, 29):
        assert lib.slc_getitem(ra, off) == complex(float(a.reshape(-1)[off]), 0.0)
    with pytest.raises(IndexError):
        lib.slc_getitem(rz, 63)
    s = synth.shifted_scene(seed=5, size=128, grid=(2, 2))
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                  s.refine_factor, s.refine_margin, s.zoom_factor)
    worker = lib.sequential(s.pairs, (s.chip, s.chip), (s.window, s.window), s.refine_margin,
                            s.refine_factor, s.zoom_factor)
    r = lib.raster(s.reference, *s.reference.shape)
    t = lib.raster(s.target, *s.target.shape)
    lib.addPairs(worker, r, t, 0, s.ref_begin.tobytes(), s.tgt_begin.tobytes())
    offsets = np.array(lib.adjust(worker), np.float32)
    assert lib.refine(worker) is None
    assert np.array_equal(offsets, hip["offsets"])


def test_tile_registration_errors(lib):
    # the host checks that keep every gather inside its raster (no kernel runs
    # on a rejected pair): slices past the raster, wrong shapes, pair indices
    # out of range, mixed pixel types, and an adjust with a pair missing
    rng = np.random.default_rng(21)
    z = (rng.standard_normal((96, 80)) + 1j * rng.standard_normal((96, 80))).astype(np.complex64)
    a = np.zeros((96, 80), np.float32)
    rz = lib.raster(z, 96, 80)
    ra = lib.raster(a, 96, 80, dtype="float32")
    worker = lib.sequential(2, (16, 16), (32, 32), 4, 2, 4)
    with pytest.raises(ValueError, match="beyond the raster"):
        lib.addReference(worker, rz, 0, (90, 0), (106, 16))
    with pytest.raises(ValueError, match="beyond the raster"):
        lib.addTarget(worker, rz, 0, (0, 60), (32, 92))
    with pytest.raises(ValueError, match="shape"):
        lib.addReference(worker, rz, 0, (0, 0), (16, 17))
    with pytest.raises(ValueError, match="out of range"):
        lib.addTarget(worker, rz, 2, (0, 0), (32, 32))
    with pytest.raises(ValueError, match="negative"):
        lib.addPairs(worker, rz, rz, 0, np.array([[0, -1]], np.int64).tobytes(),
                     np.array([[0, 0]], np.int64).tobytes())
    lib.addReference(worker, rz, 0, (0, 0), (16, 16))
    with pytest.raises(ValueError, match="pixel type"):
        lib.addTarget(worker, ra, 0, (0, 0), (32, 32))
    lib.addTarget(worker, rz, 0, (0, 0), (32, 32))
    with pytest.raises(lib.Error, match="pair 1"):
        lib.adjust(worker)
    # once every pair is registered the same worker runs
    lib.addReference(worker, rz, 1, (40, 40), (56, 56))
    lib.addTarget(worker, rz, 1, (32, 32), (64, 64))
    offsets = np.array(lib.adjust(worker), np.float32)
    assert offsets.shape == (2, 2) and np.all(np.isfinite(offsets))


@pytest.mark.parametrize("rf,rm,zf", [(3, 4, 4), (4, 3, 2)])
def test_refine_factors_other_than_two(lib, rf, rm, zf):
    # the dense U x U^T upsampler (k_upsample) for refine factors other than 2
    s = synth.shifted_scene(seed=23, size=256, grid=(3, 3), chip=32, margin=16, refine_factor=rf,
                            refine_margin=rm, zoom_factor=zf, shift=(2.4, -1.1))
    print(f"refine factor {rf}, margin {rm}, zoom {zf}: {s.pairs} pairs", flush=True)
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window, rf, rm,
                  zf)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)
    for precision in ("bf16", "fp16"):
        mixed = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                        rf, rm, zf, precision=precision)
        assert np.max(np.abs(mixed["offsets"] - exact["offsets"])) <= 1e-2
        assert np.array_equal(mixed["coarse_loc"], exact["coarse_loc"])


def test_refine_factor_one_and_a_chip_of_160(lib):
    # no upsampling (refine factor 1) and a chip above 128 samples: the mixed
    # forms' coarse stage then gathers FP32 amplitudes and quantizes them in a
    # separate pass (k_gather_detect + k_prep_mixed)
    s = synth.shifted_scene(seed=29, size=512, grid=(2, 2), chip=160, margin=16, refine_factor=1,
                            refine_margin=4, zoom_factor=8, shift=(-3.6, 2.2))
    print(f"chip 160, refine factor 1: {s.pairs} pairs", flush=True)
    hip = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window, 1, 4, 8)
    twin, exact = oracles(s)
    check_against_oracles(hip, twin, exact)
    for precision in ("bf16", "fp16"):
        mixed = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                        1, 4, 8, precision=precision)
        assert np.max(np.abs(mixed["offsets"] - exact["offsets"])) <= 1e-2
        assert np.array_equal(mixed["coarse_loc"], exact["coarse_loc"])
