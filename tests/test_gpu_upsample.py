"""The HIP upsamplers against the golden vectors (tests/golden/upsample_*.npz,
NumPy FFT zero-padding with the Nyquist bin split).

stage_upsample runs a plan's target upsampling alone on given tiles, so each
fixture goes through the kernel its size and the plan's form select:
- k_upsample2 (FP32 matrix cores, three circulant products): FP32 form, x2,
  tiles up to 96;
- k_up2_cols + k_up2_rows (banded two-pass, FP32): tiles above 96, in the FP32
  form and in the quantized modes of the BF16/FP16 forms;
- k_upsample (dense U x U^T): refine factors other than 2;
- k_upsample2h (FP16 matrix cores): the BF16/FP16 forms, complex tiles up to
  96 and real tiles up to 160.
The bar is on |y|, relative to the tile's largest |y|: 1e-5 for the FP32
paths, 2e-3 for the FP16 one (its operands carry 11 bits).
"""

import glob
import os

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def geometry(N, rf):
    """A plan whose expanded target window E = chip + 2 rm is N."""
    rm = 2 if N <= 16 else 8
    chip = N - 2 * rm
    return chip, rm


def upsample(lib, tiles, rf, precision, dtype="complex64"):
    n, N, _ = tiles.shape
    chip, rm = geometry(N, rf)
    worker = lib.sequential(n, (chip, chip), (N, N), rm, rf, 2, precision=precision)
    raw = lib.stage_upsample(worker, np.ascontiguousarray(tiles).tobytes(), n, dtype=dtype)
    return np.frombuffer(raw, np.float32).reshape(n, rf * N, rf * N)


def forms(N, rf):
    chip, _ = geometry(N, rf)
    out = [("fp32", 1e-5)]
    if chip % 16 == 0 and chip * rf <= 256:
        # (N <= 96: the FP16 matrix-core upsampler; above: the banded one, quantized)
        tol = 2e-3 if N <= 96 else 1e-5
        out += [("bf16", tol), ("fp16", tol), ("bf16x3", 1e-5)]
    return out


CASES = [(os.path.basename(p), prec, tol)
         for p in sorted(glob.glob(os.path.join(GOLDEN, "upsample_*.npz")))
         for prec, tol in forms(int(np.load(p)["x"].shape[0]), int(np.load(p)["rf"]))]


@pytest.mark.parametrize("name,precision,tol", CASES)
def test_upsampler_against_golden(lib, name, precision, tol):
    f = np.load(os.path.join(GOLDEN, name))
    x, rf, y = f["x"], int(f["rf"]), f["y"]
    # the fixture's tile, and its transpose and conjugate as more pairs
    tiles = np.stack([x, x.T, np.conj(x)]).astype(np.complex64)
    want = np.abs(np.stack([y, y.T, np.conj(y)]))
    got = upsample(lib, tiles, rf, precision)
    err = np.max(np.abs(got - want)) / np.max(want)
    print(f"{name} {precision}: max |amp - golden| / max = {err:.2e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("N,precision,tol", [(112, "fp16", 2e-3), (112, "bf16", 2e-3),
                                            (144, "fp16", 2e-3), (144, "fp32", 1e-5)])
def test_upsampler_real_tiles(lib, N, precision, tol):
    """Real-valued tiles (amplitude rasters): the FP16 upsampler's two-plane path
    up to 160 samples, against the oracle (itself pinned to the golden vectors)."""
    rng = np.random.default_rng(N)
    tiles = np.exp(0.5 * rng.standard_normal((2, N, N))).astype(np.float32)
    want = np.abs(np.stack([pyoracle.upsample(t.astype(np.complex128), 2) for t in tiles]))
    got = upsample(lib, tiles, 2, precision, dtype="float32")
    err = np.max(np.abs(got - want)) / np.max(want)
    print(f"real {N} {precision}: max |amp - oracle| / max = {err:.2e}", flush=True)
    assert err <= tol
