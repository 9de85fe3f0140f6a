"""The algorithmic work bench.py divides by each stage's time (CPU)."""

import types

import numpy as np

import bench


def scene(ref_begin, tgt_begin, shape=(512, 512)):
    raster = np.zeros(shape, np.complex64)
    return types.SimpleNamespace(pairs=len(ref_begin), chip=64, window=128, refine_factor=2,
                                 refine_margin=8, reference=raster, target=raster,
                                 ref_begin=np.array(ref_begin), tgt_begin=np.array(tgt_begin))


SCENE = scene([(0, 0)], [(0, 0)])


def test_contraction_flops_per_pair():
    # 2 C^2 P^2 (coarse) and 2 (rf C)^2 R^2 (refined): 34.61 + 35.68 MFLOP at chip 64 / window 128
    w = bench.stage_work("bf16", SCENE, "complex64")
    assert w["correlate"] == ("mfma", 2.0 * 64 * 64 * 65 * 65, "FLOP")
    assert w["correlate_refined"] == ("mfma", 2.0 * 128 * 128 * 33 * 33, "FLOP")
    assert round(w["correlate"][1] / 1e6, 2) == 34.61
    assert round(w["correlate_refined"][1] / 1e6, 2) == 35.68
    assert bench.stage_work("fp32", SCENE, "complex64")["correlate"][0] == "valu"


def test_hbm_bytes_per_pair():
    f = bench.stage_work("fp32", SCENE, "complex64")
    # gather + detect: complex64 chip and window read, FP32 amplitudes written
    assert f["gather_detect"][1] == 163840 + 81920
    # argmax reads the coarse surface once
    assert f["coarse_peak"][1] == 4 * 65 * 65
    # migrate + upsample of the 80 x 80 window: read once, FP32 160 x 160 amplitudes written
    assert f["upsample_tgt"][1] == 8 * 80 * 80 + 4 * 160 * 160
    b = bench.stage_work("bf16", SCENE, "complex64")
    # the mixed gather writes two-byte operands; FP8 one byte
    assert b["gather_prep"][1] == 8 * (64 * 64 + 128 * 128) + 2 * (64 * 64 + 128 * 128)
    assert bench.stage_work("fp8", SCENE, "complex64")["gather_prep"][1] == \
        8 * (64 * 64 + 128 * 128) + 1 * (64 * 64 + 128 * 128)
    # a float32 raster reads half the bytes
    assert bench.stage_work("fp32", SCENE, "float32")["gather_detect"][1] == 8 * (64 * 64 + 128 * 128)


def test_stage_roofline_fraction():
    st = bench.stage_roofline("bf16", SCENE, "complex64", {"correlate": 1e-3, "zoom": 1e-3})
    # 34.6 MFLOP in 1 us = 34.6 TFLOP/s of the 2.5 PFLOP/s BF16 peak
    assert st["correlate"]["unit"] == "TFLOP/s" and abs(st["correlate"]["frac"] - 0.0138) < 1e-4
    assert st["zoom"]["unit"] == "GB/s" and st["zoom"]["peak"] == 8000.0
    fp8 = bench.stage_roofline("fp8", SCENE, "complex64", {"correlate": 1e-3, "correlate_refined": 1e-3})
    assert fp8["correlate"]["peak"] == 5000.0 and fp8["correlate_refined"]["peak"] == 2500.0


def test_gathers_count_covered_pixels_once():
    # two pairs whose windows overlap by half: 128 x 192 target pixels, not 2 x 128^2;
    # the chips do not overlap
    two = scene([(0, 0), (0, 64)], [(0, 0), (0, 64)])
    g = bench.stage_work("bf16", two, "complex64")["gather_prep"][1]
    assert g == 8 * (2 * 64 * 64 + 128 * 192) + 2 * 2 * (64 * 64 + 128 * 128)
