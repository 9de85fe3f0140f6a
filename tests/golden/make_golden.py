"""Writes the golden fixtures of tests/golden/ (run from the repository root).

Vectors computed independently of the oracle, with NumPy's FFT:
  upsample_*.npz   x2 / x3 band-limited upsampling of random complex tiles
                   (zero-padded spectrum, Nyquist bin split for even sizes),
                   including 48 x 48 and 112 x 112 tiles for the GPU upsamplers
  zoom_*.npz       band-limited zoom of random real odd-size surfaces
and end-to-end cases whose expected result is fixed by construction:
  shift_e2e.npz    a 256 x 256 speckle pair related by a periodic Fourier shift
                   of (+1.375, -2.625) px plus 25 dB noise; the offsets of every
                   pair must come out near that shift.
  shift_offgrid.npz  the same construction with a shift of (+3.31, -1.69) px,
                   which is not a multiple of the zoomed cell 1/(rf zf) = 1/8 px:
                   the sub-pixel parabola moves each offset by about 0.06 px
                   there, so leaving it out (or flipping it) misses the shift.

    python tests/golden/make_golden.py [name ...]   (default: all)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ampcor_amd import synth  # noqa: E402


def fft_upsample(x, rf):
    N = x.shape[0]
    M = rf * N
    X = np.fft.fft2(x)
    w = np.ones(N)
    k = np.fft.fftfreq(N, 1.0 / N).astype(int)
    if N % 2 == 0:
        w[N // 2] = 0.5
    Y = np.zeros((M, M), dtype=complex)
    for a in range(N):
        for b in range(N):
            ka, kb = k[a], k[b]
            v = X[a, b] * w[a] * w[b]
            pa = [ka] if not (N % 2 == 0 and a == N // 2) else [N // 2, -N // 2]
            pb = [kb] if not (N % 2 == 0 and b == N // 2) else [N // 2, -N // 2]
            for ia in pa:
                for ib in pb:
                    Y[ia % M, ib % M] += v
    return np.fft.ifft2(Y) * rf * rf


def fft_zoom(g, zf):
    R = g.shape[0]
    Z = zf * R
    G = np.fft.fft2(g)
    Y = np.zeros((Z, Z), dtype=complex)
    h = (R - 1) // 2
    for a in range(-h, h + 1):
        for b in range(-h, h + 1):
            Y[a % Z, b % Z] = G[a % R, b % R]
    return np.real(np.fft.ifft2(Y)) * zf * zf


def shift_fixture(name, seed, shift):
    s = synth.shifted_scene(seed=seed, size=256, grid=(4, 4), shift=shift, snr_db=25.0)
    np.savez(os.path.join(HERE, name), reference=s.reference, target=s.target,
             ref_begin=s.ref_begin, tgt_begin=s.tgt_begin,
             geometry=np.array([s.chip, s.window, s.refine_factor, s.refine_margin, s.zoom_factor]),
             shift=np.array(shift))


def main(names=()):
    want = lambda name: not names or name in names
    rng = np.random.default_rng(20261203)
    for N, rf in [(8, 2), (10, 2), (9, 2), (6, 3)]:
        x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
        if want(f"upsample_{N}x{rf}.npz"):
            np.savez(os.path.join(HERE, f"upsample_{N}x{rf}.npz"), x=x, rf=rf, y=fft_upsample(x, rf))
    for R, zf in [(9, 4), (11, 3)]:
        g = rng.standard_normal((R, R))
        if want(f"zoom_{R}x{zf}.npz"):
            np.savez(os.path.join(HERE, f"zoom_{R}x{zf}.npz"), g=g, zf=zf, z=fft_zoom(g, zf))
    # tiles the GPU upsamplers take as target windows: 48 (chip 32 + 2 x 8) runs
    # the FP16 matrix-core upsampler of the mixed forms, 112 (chip 96 + 2 x 8)
    # the banded two-pass one of the FP32 form
    big = np.random.default_rng(20261204)
    for N in (48, 112):
        x = big.standard_normal((N, N)) + 1j * big.standard_normal((N, N))
        x = x.astype(np.complex64).astype(np.complex128)   # exactly a complex64 raster's tile
        if want(f"upsample_{N}x2.npz"):
            np.savez(os.path.join(HERE, f"upsample_{N}x2.npz"), x=x, rf=2, y=fft_upsample(x, 2))
    if want("shift_e2e.npz"):
        shift_fixture("shift_e2e.npz", 21, (1.375, -2.625))
    if want("shift_offgrid.npz"):
        shift_fixture("shift_offgrid.npz", 22, (3.31, -1.69))


if __name__ == "__main__":
    main(tuple(sys.argv[1:]))
