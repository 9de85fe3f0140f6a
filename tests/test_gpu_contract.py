"""The mixed contraction's tilings on exact integer operands.

stage_contract runs one contraction of a BF16 / FP16 plan alone.  The operands
here are small integers (exact in bfloat16 and binary16) whose products and
partial sums are integers below 2^24, so every tiling -- whatever order it
accumulates in -- must give the integer correlation
    N[u][v] = sum_ij window[u + i][v + j] chip[i][j]
exactly.  A wrong index, a dropped or doubled chip row or target column, a
misplaced output row (the packed form's interleaved tiles and its packed last
row) or a missing edge column fails bitwise.

Forms: 1 = 16 x 16 tiles of consecutive placements (chip-row ranges summed in
order), 2 = the streaming form, 3 = the packed form (surfaces of 32 n + 1
placements per axis; chips of 16 .. 160 samples and of 256, the refined chip
of the 128-sample chips), 4 / 5 = the packed form in two / four chip-row
ranges (each range's window rows start at its first chip row).  The geometries cover the scene's coarse (chip 64,
window 128) and refined (128 / 160) stages, the serving stream's chips 32 and
128 with margins 16, 32 and 64, and the packed form's multi-wave layouts (two
and four u-blocks, two v-groups).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMS = {1: "tiles", 2: "streaming", 3: "packed", 4: "packed, 2 ranges", 5: "packed, 4 ranges"}


def bits(x, precision):
    """16-bit patterns of integer-valued float32 x (exact in either format)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if precision == "fp16":
        return x.astype(np.float16).view(np.uint16)
    return (x.view(np.uint32) >> 16).astype(np.uint16)


def brute(chip, window):
    """Integer correlation, one pair at a time (int64, exact)."""
    n, c, _ = chip.shape
    w = window.shape[1]
    p = w - c + 1
    out = np.zeros((n, p, p), np.int64)
    cw = chip.astype(np.int64)
    ww = window.astype(np.int64)
    for i in range(c):
        for j in range(c):
            out += cw[:, i, j, None, None] * ww[:, i:i + p, j:j + p]
    return out


# (chip, window, refine margin): coarse stage C x W, refined stage 2C x 2(C + 2 rm)
GEOMETRIES = [(64, 128, 8), (32, 64, 8), (32, 96, 16), (32, 160, 8), (64, 192, 8),
              (128, 192, 8), (16, 48, 8)]


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
@pytest.mark.parametrize("chip,window,rm", GEOMETRIES)
def test_tilings_exact_on_integers(lib, chip, window, rm, precision):
    n = 5
    rng = np.random.default_rng(chip * 1000 + window + rm)
    worker = lib.sequential(n, (chip, chip), (window, window), rm, 2, 8, precision=precision)
    ran = 0
    for stage, (c, w) in enumerate([(chip, window), (2 * chip, 2 * (chip + 2 * rm))]):
        # |sum| <= c^2 v^2 < 2^24
        v = 16 if c <= 128 else 8
        cq = rng.integers(-v, v + 1, size=(n, c, c)).astype(np.float32)
        tq = rng.integers(-v, v + 1, size=(n, w, w)).astype(np.float32)
        # zero-mean rows and a constant block: the forms must not assume either away
        tq[0, : w // 2, : w // 3] = v
        want = brute(cq, tq)
        assert np.abs(want).max() < 2 ** 24
        p = w - c + 1
        for form, name in FORMS.items():
            try:
                raw = lib.stage_contract(worker, stage, form, bits(cq, precision).tobytes(), c,
                                         bits(tq, precision).tobytes(), w, n)
            except Exception as e:  # the form does not apply to this geometry
                assert "does not apply" in str(e) or "exceeds" in str(e), str(e)
                if form >= 3:   # built for chips of 16 .. 160 samples and 256
                    ranges = 1 << (form - 3)
                    assert ((p - 1) % 32 or c % 16 or (c > 160 and c != 256)
                            or (c // ranges) % 4), (stage, p, c, ranges)
                continue
            got = np.frombuffer(raw, np.float32).reshape(n, p, p)
            bad = np.argwhere(got.astype(np.int64) != want)
            print(f"{precision} stage {stage} c {c} w {w} p {p} {name}: "
                  f"{len(bad)} of {got.size} differ", flush=True)
            assert len(bad) == 0, (name, stage, bad[:8].tolist())
            ran += 1
    assert ran >= 2


def test_packed_form_is_the_default_where_it_applies(lib):
    """The scene's plans (chip 64 / window 128, refine margin 8: P = 65, R = 33)
    run the packed form in both stages: form 0 (the plan's tiling) equals form 3."""
    n = 3
    rng = np.random.default_rng(5)
    worker = lib.sequential(n, (64, 64), (128, 128), 8, 2, 8, precision="bf16")
    for stage, (c, w) in enumerate([(64, 128), (128, 160)]):
        cq = bits(rng.integers(-8, 9, size=(n, c, c)), "bf16").tobytes()
        tq = bits(rng.integers(-8, 9, size=(n, w, w)), "bf16").tobytes()
        cur = lib.stage_contract(worker, stage, 0, cq, c, tq, w, n)
        packed = lib.stage_contract(worker, stage, 3, cq, c, tq, w, n)
        assert cur == packed


def test_autotune_large_geometry(lib):
    """Chip 128 / window 256 (a 129 x 129 surface: four u-blocks, several
    v-groups) in FP16: every tiling the autotuner launches must fit its
    kernel's workgroup limit, and the tuned plan gives the same offsets to
    the forms' accumulation-order tolerance."""
    from ampcor_amd import synth

    s = synth.amplitude_scene(seed=3, size=1024, grid=(4, 4), max_shift=2.0)
    worker = lib.sequential(s.pairs, (s.chip, s.chip), (s.window, s.window), s.refine_margin,
                            s.refine_factor, s.zoom_factor, precision="fp16")
    ref = lib.raster(s.reference, *s.reference.shape, dtype="float32")
    tgt = lib.raster(s.target, *s.target.shape, dtype="float32")
    lib.addPairs(worker, ref, tgt, 0, s.ref_begin.tobytes(), s.tgt_begin.tobytes())
    before = np.array(lib.adjust(worker), dtype=np.float64)
    lib.autotune(worker, 1)
    after = np.array(lib.adjust(worker), dtype=np.float64)
    print(f"{s.pairs} pairs: max |tuned - default| = {np.abs(after - before).max():.2e} px",
          flush=True)
    assert np.abs(after - before).max() < 1e-3
