"""Large-chip amplitude case (chip 128, margin 64, zoom 16, float32 rasters) on
a 2048 x 2048 cut of the scene geometry (256 pairs), FP32 and FP16 forms vs
the oracles.

The FP32 bars are the drop-in's: coarse surfaces bitwise equal to the FP32
twin, offsets within 1e-3 px of the FP64 oracle, coarse peak, refined peak and
SNR within 1e-5 relative.  The background of a 129 x 129 surface over
textured amplitudes (a large mean under a small variance) is where a single
FP32 accumulation chain over all 16384 products loses precision; a cut with
only a handful of pairs would not see it.
"""

import numpy as np
import pytest

from ampcor_amd import synth
from oracle import pyoracle
from tests.gpu_helpers import run_hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    # the field keeps the scene gradient: 12 px over 8192 lines is 3 px over this 2048 cut
    s = synth.amplitude_scene(seed=14, size=2048, grid=(18, 18), max_shift=3.0)
    print(f"amplitude scene: {s.pairs} pairs", flush=True)
    assert s.pairs >= 200
    return s


@pytest.fixture(scope="module")
def oracles(scene):
    s = scene
    args = (s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window, s.refine_factor,
            s.refine_margin, s.zoom_factor)
    return (pyoracle.adjust(*args, form="fp32", surfaces=True),
            pyoracle.adjust(*args, form="fp64", surfaces=True))


def run(lib, s, precision):
    return run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                   s.refine_factor, s.refine_margin, s.zoom_factor, precision=precision)


def rel_err(x, y):
    return float(np.max(np.abs(x.astype(np.float64) - y) / np.abs(y)))


def test_large_chip_fp32(lib, scene, oracles):
    twin, exact = oracles
    hip = run(lib, scene, "fp32")
    assert np.array_equal(hip["gamma"].view(np.uint32), twin["gamma"].view(np.uint32))
    assert np.array_equal(hip["coarse_loc"], twin["coarse_loc"])
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])
    errs = {k: rel_err(hip[k], exact[k]) for k in ("coarse_peak", "refined_peak", "snr")}
    off = float(np.max(np.abs(hip["offsets"] - exact["offsets"])))
    print(f"fp32 large chip, {scene.pairs} pairs: max |offset - fp64| = {off:.2e} px, "
          f"max relative error vs fp64: {errs}", flush=True)
    assert off <= 1e-3
    for k, e in errs.items():
        assert e <= 1e-5, (k, e)
    assert np.max(np.abs(hip["offsets"] - scene.truth)) < 0.1


def test_large_chip_fp16(lib, scene, oracles):
    _, exact = oracles
    hip = run(lib, scene, "fp16")
    err = np.abs(hip["offsets"] - exact["offsets"])
    print(f"fp16 large chip: max |offset - fp64| = {err.max():.2e} px", flush=True)
    assert err.max() <= 1e-2
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])


def test_large_chip_fp8(lib, scene, oracles):
    # FP8 coarse surface (e4m3 operands, FP32 accumulate), FP16 refinement.
    # The peak guard re-scores the near-ties of the e4m3 surface in FP64, so
    # every coarse peak location is the FP64 oracle's
    _, exact = oracles
    hip = run(lib, scene, "fp8")
    err = np.abs(hip["offsets"] - exact["offsets"])
    g = hip["guard"]
    print(f"fp8 large chip: max |offset - fp64| = {err.max():.2e} px, guard {g}", flush=True)
    assert err.max() <= 1e-2
    assert g["overflow"] == 0
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])
    assert np.array_equal(hip["window_loc"], exact["window_loc"])
