"""The FP8 form: coarse contraction on OCP e4m3 matrix cores
(v_mfma_f32_16x16x128_f8f6f4, FP32 accumulate), refined re-correlation as the
FP16 form, FP64 window statistics.

Bars:
- the coarse surfaces equal the FP64 normalized correlation of exactly the
  e4m3 operands (restated in NumPy, tests/fp8_ref.py) to 2e-3: this pins the
  contraction, the quantizer and the normalizer of the FP8 stage;
- the coarse peak lands on the FP64 oracle's placement (per-tile mean removal
  and scaling keep the peak), and the refined offsets stay within 0.01 px of
  the FP64 oracle.
"""

import numpy as np
import pytest

from ampcor_amd import synth
from oracle import pyoracle
from tests.fp8_ref import quantize_pair
from tests.gpu_helpers import run_hip

pytestmark = pytest.mark.gpu

OFFSET_TOL = 0.01      # px vs the FP64 oracle
GAMMA_TOL = 2e-3       # coarse gamma vs the FP64 NCC of the e4m3 operands


def fp64(s):
    return pyoracle.adjust(s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                           s.refine_factor, s.refine_margin, s.zoom_factor, form="fp64")


def run(lib, s, precision="fp8"):
    return run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                   s.refine_factor, s.refine_margin, s.zoom_factor, precision=precision)


def test_fp8_coarse_surface_is_the_ncc_of_the_e4m3_operands(lib):
    s = synth.shifted_scene()
    hip = run(lib, s)
    amp_r, amp_t = np.abs(s.reference), np.abs(s.target)
    C, W = s.chip, s.window
    worst = 0.0
    for k in range(0, s.pairs, 7):
        (r0, c0), (t0, u0) = s.ref_begin[k], s.tgt_begin[k]
        rq, tq = quantize_pair(amp_r[r0:r0 + C, c0:c0 + C], amp_t[t0:t0 + W, u0:u0 + W])
        g = pyoracle.ncc64(rq, tq)
        worst = max(worst, float(np.abs(hip["gamma"][k] - g).max()))
    print(f"max |gamma_fp8 - ncc64(e4m3 operands)| = {worst:.2e}", flush=True)
    assert worst <= GAMMA_TOL


def test_fp8_shifted_scene(lib):
    s = synth.shifted_scene()
    hip, exact = run(lib, s), fp64(s)
    err = np.abs(hip["offsets"] - exact["offsets"])
    print(f"fp8: max |offset - fp64| = {err.max():.2e} px", flush=True)
    assert err.max() <= OFFSET_TOL
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])


def test_fp8_field_scene(lib):
    s = synth.field_scene(seed=6, lines=1024, samples=1024, grid=(12, 12), range_ramp=3.0,
                          azimuth_quadratic=1.5)
    hip, exact = run(lib, s), fp64(s)
    err = np.abs(hip["offsets"] - exact["offsets"])
    print(f"fp8: {s.pairs} pairs, max |offset - fp64| = {err.max():.2e} px, peak guard "
          f"{hip['guard']}", flush=True)
    assert err.max() <= OFFSET_TOL
    # the peak guard keeps every coarse peak where FP64 puts it
    assert np.array_equal(hip["coarse_loc"], exact["coarse_loc"])


def test_fp8_planted_chip(lib):
    ref, tgt, rb, tb = synth.planted(32, 8, seed=4)
    hip = run_hip(lib, ref, tgt, rb, tb, 32, 48, 2, 4, 4, precision="fp8")
    P = 17
    for i in range(P):
        for j in range(P):
            k = i * P + j
            assert tuple(hip["coarse_loc"][k]) == (i, j)
            assert abs(hip["gamma"][k, i, j] - 1.0) < 2e-2


def test_fp8_float32_rasters(lib):
    rng = np.random.default_rng(12)
    ref = synth.lognormal_texture((384, 384), 0.5, rng)
    tgt = np.real(synth.fourier_shift(ref, -2.5, 1.75)).astype(np.float32)
    tgt = tgt * (1 + 0.05 * rng.standard_normal(tgt.shape)).astype(np.float32)
    s = synth.Scene("amp", ref, tgt, 32, 16, 2, 8, 8, (4, 4))
    s.plan()
    hip, exact = run(lib, s), fp64(s)
    assert np.abs(hip["offsets"] - exact["offsets"]).max() <= OFFSET_TOL


def test_fp8_rejects_unsupported_geometry(lib):
    # the FP8 coarse stage needs the one-pass gather (chip <= 128)
    with pytest.raises(Exception):
        lib.sequential(4, (144, 144), (176, 176), 8, 2, 4, precision="fp8")
