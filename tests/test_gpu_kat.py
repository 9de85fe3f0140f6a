"""Known-answer tests of the reference's own test suite, run against the HIP
kernels (the same cases pin the oracle in tests/test_oracle.py):

- SumArea closed forms (tests/libampcor/sumarea.cc:44-69,
  tests/libampcor/sumarea-sum.cc:47-148) and the GPU SAT against them
  (tests/cuda/tile-sat.cu:216-270): the FP32 form's FP64 window sums of windows
  of ones, of an integer ramp and of random integers equal the exact integer
  box sums bitwise, in both stages' geometries;
- the tile-correlate closed form (tests/cuda/tile-correlate.cu:359-388): a chip
  of ones against ones padded by zeros gives (C - |m - r|)(C - |m - c|), through
  every tiling of the matrix-core contraction;
- detection (tests/libampcor_cuda/detect.cc:202-215): the amplitudes the FP32
  form detects equal |z| rounded once to float32 (exact on Pythagorean triples),
  and |x| for real rasters;
- migration (tests/libampcor_cuda/migrate.cc:350-426): the refined tiles hold
  the pixels of the nudged windows -- with refine factor 1 the refined
  amplitudes are those windows' |z| bitwise, and with refine factor 2 their
  even samples are, to the 1 ulp of the upsampler's FP32 square root.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHIP, WINDOW, RM, RF, ZF = 32, 64, 8, 2, 4


def box_sums(t, c):
    """Exact integer sums of t and t^2 over every c x c placement (int64)."""
    t = t.astype(np.int64)
    n, w, _ = t.shape
    p = w - c + 1
    s1 = np.zeros((n, p, p), np.int64)
    s2 = np.zeros((n, p, p), np.int64)
    for i in range(c):
        for j in range(c):
            blk = t[:, i:i + p, j:j + p]
            s1 += blk
            s2 += blk * blk
    return s1, s2


@pytest.mark.parametrize("stage", [0, 1])
def test_window_sums_are_the_exact_box_sums(lib, stage):
    n = 3
    worker = lib.sequential(n, (CHIP, CHIP), (WINDOW, WINDOW), RM, RF, ZF)
    c, w = (CHIP, WINDOW) if stage == 0 else (RF * CHIP, RF * (CHIP + 2 * RM))
    rng = np.random.default_rng(11 + stage)
    i, j = np.meshgrid(np.arange(w), np.arange(w), indexing="ij")
    windows = np.stack([np.ones((w, w)),                       # the area table
                        i + 2 * j + 1,                         # a ramp
                        rng.integers(-50, 51, size=(w, w))]).astype(np.float32)
    raw = lib.stage_window_sums(worker, stage, windows.tobytes(), w, n)
    p = w - c + 1
    got = np.frombuffer(raw, np.float64).reshape(n, p, p, 2)
    s1, s2 = box_sums(windows, c)
    assert np.all(s1[0] == c * c) and np.all(s2[0] == c * c)
    print(f"stage {stage}: window {w}, chip {c}, {p} x {p} placements", flush=True)
    assert np.array_equal(got[..., 0], s1.astype(np.float64))
    assert np.array_equal(got[..., 1], s2.astype(np.float64))


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_tile_correlate_closed_form(lib, precision):
    C, m = CHIP, (WINDOW - CHIP) // 2
    worker = lib.sequential(1, (C, C), (WINDOW, WINDOW), RM, RF, ZF, precision=precision)
    one = np.float32(1.0).view(np.uint32) >> 16 if precision == "bf16" else \
        np.float16(1.0).view(np.uint16)
    chip = np.full((1, C, C), one, np.uint16)
    tgt = np.zeros((1, WINDOW, WINDOW), np.uint16)
    tgt[0, m:m + C, m:m + C] = one
    r, c = np.meshgrid(np.arange(2 * m + 1), np.arange(2 * m + 1), indexing="ij")
    want = ((C - np.abs(m - r)) * (C - np.abs(m - c))).astype(np.float32)
    ran = 0
    for form in (1, 2, 3):
        try:
            raw = lib.stage_contract(worker, 0, form, chip.tobytes(), C, tgt.tobytes(), WINDOW, 1)
        except ValueError as e:   # the form does not apply to this geometry
            assert "does not apply" in str(e), str(e)
            continue
        got = np.frombuffer(raw, np.float32).reshape(2 * m + 1, 2 * m + 1)
        assert np.array_equal(got, want), form
        ran += 1
    assert ran >= 2


@pytest.mark.parametrize("real", [False, True])
def test_detected_amplitudes(lib, real):
    rng = np.random.default_rng(21)
    lines = samples = 2 * WINDOW
    if real:
        raster = (rng.standard_normal((lines, samples)) * 8).astype(np.float32)
        want_of = np.abs
    else:
        raster = (rng.standard_normal((lines, samples)) +
                  1j * rng.standard_normal((lines, samples))).astype(np.complex64)
        # Pythagorean triples at several scales: |z| exact
        for k, (a, b) in enumerate([(3, 4), (5, 12), (8, 15), (7, 24), (20, 21)]):
            raster[k, :8] = [complex(a * s, b * s) for s in (1, 2, 0.5, 1024, 2 ** -10, 3, 5, 7)]
        raster[6, :4] = [0, 1j, -1, -3 - 4j]

        def want_of(z):
            re = z.real.astype(np.float64)
            im = z.imag.astype(np.float64)
            return np.sqrt(re * re + im * im).astype(np.float32)
    ref_begin = np.array([[0, 0], [40, 24]], np.int64)
    tgt_begin = np.array([[0, 0], [24, 40]], np.int64)
    n = len(ref_begin)
    worker = lib.sequential(n, (CHIP, CHIP), (WINDOW, WINDOW), RM, RF, ZF)
    ras = lib.raster(raster, lines, samples, dtype="float32" if real else "complex64")
    lib.addPairs(worker, ras, ras, 0, ref_begin.tobytes(), tgt_begin.tobytes())
    lib.adjust(worker, offsets=False)
    chips, wins = lib.stage_amplitudes(worker, 0, CHIP, WINDOW)
    chips = np.frombuffer(chips, np.float32).reshape(n, CHIP, CHIP)
    wins = np.frombuffer(wins, np.float32).reshape(n, WINDOW, WINDOW)
    for k in range(n):
        r0, c0 = ref_begin[k]
        t0, s0 = tgt_begin[k]
        assert np.array_equal(chips[k], want_of(raster[r0:r0 + CHIP, c0:c0 + CHIP]))
        assert np.array_equal(wins[k], want_of(raster[t0:t0 + WINDOW, s0:s0 + WINDOW]))
    if not real:
        assert chips[0][0, 0] == 5 and chips[0][1, 1] == 26 and chips[0][6, 3] == 5


def test_amplitudes_are_kept_by_fp32_plans_only(lib):
    worker = lib.sequential(1, (CHIP, CHIP), (WINDOW, WINDOW), RM, RF, ZF, precision="bf16")
    with pytest.raises(ValueError, match="FP32"):
        lib.stage_amplitudes(worker, 0, CHIP, WINDOW)


@pytest.mark.parametrize("rf", [1, 2])
def test_refined_tiles_hold_the_migrated_windows(lib, rf):
    rng = np.random.default_rng(31 + rf)
    lines = samples = 2 * WINDOW
    raster = (rng.standard_normal((lines, samples)) +
              1j * rng.standard_normal((lines, samples))).astype(np.complex64)
    ref_begin = np.array([[8, 16], [40, 24], [60, 60]], np.int64)
    tgt_begin = ref_begin - (WINDOW - CHIP) // 2
    tgt_begin = np.clip(tgt_begin, 0, lines - WINDOW)
    n = len(ref_begin)
    worker = lib.sequential(n, (CHIP, CHIP), (WINDOW, WINDOW), RM, rf, ZF)
    ras = lib.raster(raster, lines, samples)
    lib.addPairs(worker, ras, ras, 0, ref_begin.tobytes(), tgt_begin.tobytes())
    lib.adjust(worker, offsets=False)
    E = CHIP + 2 * RM
    chips, wins = lib.stage_amplitudes(worker, 1, rf * CHIP, rf * E)
    chips = np.frombuffer(chips, np.float32).reshape(n, rf * CHIP, rf * CHIP)
    wins = np.frombuffer(wins, np.float32).reshape(n, rf * E, rf * E)
    top = np.frombuffer(lib.stage_locations(worker)["window"], np.int32).reshape(n, 2)

    def amp(z):
        return np.sqrt(z.real.astype(np.float64) ** 2 + z.imag.astype(np.float64) ** 2).astype(
            np.float32)
    for k in range(n):
        r0, c0 = ref_begin[k]
        t0, s0 = tgt_begin[k] + top[k]
        want_c = amp(raster[r0:r0 + CHIP, c0:c0 + CHIP])
        want_w = amp(raster[t0:t0 + E, s0:s0 + E])
        print(f"rf {rf} pair {k}: window at {tuple(top[k])}", flush=True)
        if rf == 1:
            assert np.array_equal(chips[k], want_c)
            assert np.array_equal(wins[k], want_w)
        else:   # the x2 upsampler detects with the FP32 square root (within 1 ulp)
            np.testing.assert_array_max_ulp(chips[k][::2, ::2], want_c, maxulp=1)
            np.testing.assert_array_max_ulp(wins[k][::2, ::2], want_w, maxulp=1)
