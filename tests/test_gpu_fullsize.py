"""The full-size workloads through the C ABI module, checked against the CPU
oracles on an evenly spaced sample of their pairs and, for every pair,
through properties that do not depend on the oracle's size limits.

Workloads (the sizes bench.py measures):
- the scene: 8192 x 4096 complex64 speckle, the secondary resampled by a
  smooth field (range ramp 0-5 px, azimuth quadratic up to 2 px) plus 15 dB
  noise, 120 x 60 grid: 6844 valid pairs; chip 64, window 128, refine x2
  margin 8, zoom 8;
- the large-chip case: 8192 x 8192 float32 lognormal amplitudes, the
  secondary shifted by a smooth field up to 12 px plus 10 % multiplicative
  noise, 64 x 32 grid: 1984 valid pairs; chip 128, window 256, zoom 16.

Each plan is autotuned first, as bench.py runs it.  Sample: every 16th pair
(428 and 124 pairs).  FP32 form: coarse surfaces, peak locations, nudged
windows and coarse peaks bitwise equal to the FP32 twin; offsets within
1e-3 px and coarse peak, refined peak and SNR within 1e-5 relative of the
FP64 oracle.  Mixed forms: coarse peak locations equal to the FP64
oracle's, offsets within 0.01 px of it.  Every pair: offsets finite and
within 0.1 px of the synthetic field (the FP64 oracle itself stays within
0.05 px of it on these scenes).  Batch independence: the sampled pairs run
as a plan of their own give the FP32 form's coarse surfaces bitwise and its
offsets within 1e-4 px (the refined tilings may differ).
"""

import numpy as np
import pytest

from ampcor_amd import synth
from oracle import pyoracle

pytestmark = pytest.mark.gpu

STRIDE = 16


@pytest.fixture(scope="module")
def scene():
    s = synth.field_scene()
    print(f"scene: {s.pairs} pairs", flush=True)
    assert s.pairs == 6844
    return s


@pytest.fixture(scope="module")
def large():
    s = synth.amplitude_scene()
    print(f"large-chip scene: {s.pairs} pairs", flush=True)
    assert s.pairs == 1984
    return s


def sample(s):
    return np.arange(0, s.pairs, STRIDE)


def oracle(s, idx, form):
    return pyoracle.adjust(s.reference, s.target, s.ref_begin[idx], s.tgt_begin[idx], s.chip,
                           s.window, s.refine_factor, s.refine_margin, s.zoom_factor, form=form,
                           surfaces=form == "fp32")


def run(lib, s, precision, idx=None, tune=True):
    """One plan over the scene's pairs (or the pairs idx); the products of
    every pair, and the coarse surfaces of the FP32 form."""
    rb = np.ascontiguousarray(s.ref_begin if idx is None else s.ref_begin[idx], dtype=np.int64)
    tb = np.ascontiguousarray(s.tgt_begin if idx is None else s.tgt_begin[idx], dtype=np.int64)
    n = len(rb)
    worker = lib.sequential(n, (s.chip, s.chip), (s.window, s.window), s.refine_margin,
                            s.refine_factor, s.zoom_factor, precision=precision)
    dtype = "complex64" if s.reference.dtype == np.complex64 else "float32"
    ref = lib.raster(s.reference, *s.reference.shape, dtype=dtype)
    tgt = lib.raster(s.target, *s.target.shape, dtype=dtype)
    lib.addPairs(worker, ref, tgt, 0, rb.tobytes(), tb.tobytes())
    if tune:
        lib.adjust(worker, offsets=False)
        lib.autotune(worker, 1)
    offsets = np.array(lib.adjust(worker), dtype=np.float32).reshape(n, 2)
    prods = {k: np.frombuffer(v, np.float32) for k, v in lib.products(worker).items()}
    locs = {k: np.frombuffer(v, np.int32).reshape(n, 2)
            for k, v in lib.stage_locations(worker).items()}
    out = {"tuple_offsets": offsets, "offsets": prods["offsets"].reshape(n, 2),
           "coarse_peak": prods["coarse_peak"], "refined_peak": prods["refined_peak"],
           "snr": prods["snr"], "coarse_loc": locs["coarse"], "window_loc": locs["window"]}
    if precision == "fp32":
        P = lib.dims(worker)[0]
        out["gamma"] = np.frombuffer(lib.stage_gamma(worker), np.float32).reshape(n, P, P)
    return out


def rel_err(x, y):
    return float(np.max(np.abs(x.astype(np.float64) - y) / np.abs(y)))


def check_every_pair(hip, s, name):
    err = np.abs(hip["offsets"] - s.truth)
    print(f"{name}: {len(err)} pairs, max |offset - field| = {err.max():.3f} px", flush=True)
    assert np.all(np.isfinite(hip["offsets"]))
    assert err.max() < 0.1
    assert np.array_equal(hip["tuple_offsets"], hip["offsets"])


def check_fp32(hip, s, idx, name):
    twin, exact = oracle(s, idx, "fp32"), oracle(s, idx, "fp64")
    assert np.array_equal(hip["gamma"][idx].view(np.uint32), twin["gamma"].view(np.uint32))
    for k in ("coarse_loc", "window_loc"):
        assert np.array_equal(hip[k][idx], twin[k]), k
    assert np.array_equal(hip["coarse_peak"][idx].view(np.uint32),
                          twin["coarse_peak"].view(np.uint32))
    assert np.array_equal(hip["coarse_loc"][idx], exact["coarse_loc"])
    off = float(np.max(np.abs(hip["offsets"][idx] - exact["offsets"])))
    errs = {k: rel_err(hip[k][idx], exact[k]) for k in ("coarse_peak", "refined_peak", "snr")}
    print(f"{name} fp32, {len(idx)} sampled pairs: max |offset - fp64| = {off:.2e} px, "
          f"relative errors vs fp64 {errs}", flush=True)
    assert off <= 1e-3
    for k, e in errs.items():
        assert e <= 1e-5, (k, e)


def check_mixed(hip, s, idx, name, precision):
    exact = oracle(s, idx, "fp64")
    off = float(np.max(np.abs(hip["offsets"][idx] - exact["offsets"])))
    print(f"{name} {precision}, {len(idx)} sampled pairs: max |offset - fp64| = {off:.2e} px",
          flush=True)
    assert np.array_equal(hip["coarse_loc"][idx], exact["coarse_loc"])
    assert off <= 1e-2


def check_batch_independent(lib, hip, s, idx):
    alone = run(lib, s, "fp32", idx, tune=False)
    assert np.array_equal(alone["gamma"].view(np.uint32), hip["gamma"][idx].view(np.uint32))
    assert np.array_equal(alone["coarse_loc"], hip["coarse_loc"][idx])
    assert np.max(np.abs(alone["offsets"] - hip["offsets"][idx])) <= 1e-4


def test_scene_fp32(lib, scene):
    idx = sample(scene)
    hip = run(lib, scene, "fp32")
    check_every_pair(hip, scene, "scene fp32")
    check_fp32(hip, scene, idx, "scene")
    check_batch_independent(lib, hip, scene, idx)


@pytest.mark.parametrize("precision", ["bf16", "fp16", "bf16x3", "fp8"])
def test_scene_mixed(lib, scene, precision):
    idx = sample(scene)
    hip = run(lib, scene, precision)
    check_every_pair(hip, scene, f"scene {precision}")
    check_mixed(hip, scene, idx, "scene", precision)


def test_large_fp32(lib, large):
    idx = sample(large)
    hip = run(lib, large, "fp32")
    check_every_pair(hip, large, "large fp32")
    check_fp32(hip, large, idx, "large")
    check_batch_independent(lib, hip, large, idx)


@pytest.mark.parametrize("precision", ["fp16", "bf16x3", "fp8"])
def test_large_mixed(lib, large, precision):
    idx = sample(large)
    hip = run(lib, large, precision)
    check_every_pair(hip, large, f"large {precision}")
    check_mixed(hip, large, idx, "large", precision)
