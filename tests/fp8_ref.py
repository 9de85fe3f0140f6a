"""OCP FP8 e4m3 rounding in NumPy, for checking the FP8 form's operands.

e4m3: 1 sign, 4 exponent (bias 7), 3 mantissa bits; normal binades 2^-6 ..
2^8, subnormal quantum 2^-9, largest finite 448 (no infinities). Values are
rounded to nearest, ties to even, after clamping to +-448 (as the GPU
quantizer does before converting).
"""

import numpy as np

E4M3_MAX = 448.0


def e4m3(x):
    """x (float32 array) rounded to the nearest e4m3 value, as float64."""
    x = np.clip(np.asarray(x, dtype=np.float64), -E4M3_MAX, E4M3_MAX)
    a = np.abs(x)
    _, e = np.frexp(a)            # a = m 2^e, m in [0.5, 1)
    # quantum: 2^(e - 4) in the normal range (3 mantissa bits), 2^-9 below 2^-6
    q = np.where(e - 1 >= -6, np.ldexp(1.0, e - 4), 2.0 ** -9)
    return np.sign(x) * np.round(a / q) * q


def quantize_pair(chip_amp, window_amp):
    """The FP8 operands of one pair as the GPU writes them: the chip's mean
    amplitude removed from both tiles, the chip scaled so its peak amplitude
    lands in [128, 256), the window by the scale for 16x that peak."""
    chip_amp = np.asarray(chip_amp, dtype=np.float32)
    window_amp = np.asarray(window_amp, dtype=np.float32)
    cr = np.float32(np.float64(chip_amp).mean())
    rm = float(chip_amp.max())

    def scale(v):
        return 2.0 ** (8 - np.frexp(v)[1]) if v > 0 else 1.0

    ar, at = np.float32(scale(rm)), np.float32(scale(np.float32(16.0) * np.float32(rm)))
    rq = e4m3((chip_amp - cr) * ar)
    tq = e4m3((window_amp - cr) * at)
    return rq, tq
