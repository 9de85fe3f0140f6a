"""Pins the NumPy e4m3 rounding used to check the FP8 form: the values an
MI355X's own conversion (v_cvt_pk_fp8_f32) gave for these inputs, the OCP
e4m3 code points, and round-to-nearest-even at a tie."""

import numpy as np

from tests.fp8_ref import e4m3


def test_e4m3_known_values():
    x = np.array([1.0, 0.5, 448.0, 300.0, -2.0, 2.0 ** -9, 240.0, 0.1], dtype=np.float32)
    want = np.array([1.0, 0.5, 448.0, 288.0, -2.0, 2.0 ** -9, 240.0, 0.1015625])
    np.testing.assert_array_equal(e4m3(x), want)


def test_e4m3_grid_and_ties():
    # every binade from 2^-6 to 2^8 holds 8 values 2^e (1 + k/8)
    for e in range(-6, 9):
        for k in range(8):
            v = 2.0 ** e * (1 + k / 8)
            if v <= 448.0:
                assert e4m3(np.float32(v)) == v
    # subnormals: multiples of 2^-9 below 2^-6
    for k in range(1, 8):
        assert e4m3(np.float32(k * 2.0 ** -9)) == k * 2.0 ** -9
    # ties go to the even mantissa: 1 + 1/16 -> 1, 1 + 3/16 -> 1.25
    assert e4m3(np.float32(1 + 1 / 16)) == 1.0
    assert e4m3(np.float32(1 + 3 / 16)) == 1.25
    # beyond the largest finite value: clamped
    assert e4m3(np.float32(1000.0)) == 448.0 and e4m3(np.float32(-1000.0)) == -448.0
    # tiny values round to zero (half the subnormal quantum ties to even: 0)
    assert e4m3(np.float32(2.0 ** -10)) == 0.0
