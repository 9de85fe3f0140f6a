"""Large-chip amplitude case (chip 128, margin 64, zoom 16, float32 rasters) on
a 1024 x 1024 cut of the scene geometry, FP32 and FP16 forms vs the oracles."""

import numpy as np
import pytest

from ampcor_amd import synth
from oracle import pyoracle
from tests.gpu_helpers import run_hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    # the field keeps the scene gradient: 12 px over 8192 lines is 1.5 px over this 1024 cut
    s = synth.amplitude_scene(seed=14, size=1024, grid=(6, 6), max_shift=1.5)
    print(f"amplitude scene: {s.pairs} pairs", flush=True)
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


def test_large_chip_fp32(lib, scene, oracles):
    twin, exact = oracles
    hip = run(lib, scene, "fp32")
    assert np.array_equal(hip["gamma"].view(np.uint32), twin["gamma"].view(np.uint32))
    assert np.array_equal(hip["coarse_loc"], twin["coarse_loc"])
    np.testing.assert_allclose(hip["offsets"], exact["offsets"], rtol=0, atol=1e-3)
    np.testing.assert_allclose(hip["coarse_peak"], exact["coarse_peak"], rtol=1e-5)
    np.testing.assert_allclose(hip["snr"], exact["snr"], rtol=1e-5)
    assert np.max(np.abs(hip["offsets"] - scene.truth)) < 0.1


def test_large_chip_fp16(lib, scene, oracles):
    _, exact = oracles
    hip = run(lib, scene, "fp16")
    err = np.abs(hip["offsets"] - exact["offsets"])
    print(f"fp16 large chip: max |offset - fp64| = {err.max():.2e} px", flush=True)
    assert err.max() <= 1e-2
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])


def test_large_chip_fp8(lib, scene, oracles):
    # FP8 coarse surface (e4m3 operands, FP32 accumulate), FP16 refinement:
    # the refined search re-centres a coarse peak the e4m3 rounding may move,
    # so the bar is on the offsets only
    _, exact = oracles
    hip = run(lib, scene, "fp8")
    err = np.abs(hip["offsets"] - exact["offsets"])
    moved = int(np.any(hip["coarse_loc"] != exact["coarse_loc"], axis=1).sum())
    print(f"fp8 large chip: max |offset - fp64| = {err.max():.2e} px, "
          f"{moved} coarse peaks moved", flush=True)
    assert err.max() <= 1e-2
