"""Writes the golden fixtures of tests/golden/ (run from the repository root).

Vectors computed independently of the oracle, with NumPy's FFT:
  upsample_*.npz   x2 / x3 band-limited upsampling of random complex tiles
                   (zero-padded spectrum, Nyquist bin split for even sizes)
  zoom_*.npz       band-limited zoom of random real odd-size surfaces
and one end-to-end case whose expected result is fixed by construction:
  shift_e2e.npz    a 256 x 256 speckle pair related by a periodic Fourier shift
                   of (+1.375, -2.625) px plus 25 dB noise; the offsets of every
                   pair must come out near that shift.
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


def main():
    rng = np.random.default_rng(20261203)
    for N, rf in [(8, 2), (10, 2), (9, 2), (6, 3)]:
        x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
        np.savez(os.path.join(HERE, f"upsample_{N}x{rf}.npz"), x=x, rf=rf, y=fft_upsample(x, rf))
    for R, zf in [(9, 4), (11, 3)]:
        g = rng.standard_normal((R, R))
        np.savez(os.path.join(HERE, f"zoom_{R}x{zf}.npz"), g=g, zf=zf, z=fft_zoom(g, zf))
    s = synth.shifted_scene(seed=21, size=256, grid=(4, 4), shift=(1.375, -2.625), snr_db=25.0)
    np.savez(os.path.join(HERE, "shift_e2e.npz"), reference=s.reference, target=s.target,
             ref_begin=s.ref_begin, tgt_begin=s.tgt_begin,
             geometry=np.array([s.chip, s.window, s.refine_factor, s.refine_margin, s.zoom_factor]),
             shift=np.array([1.375, -2.625]))


if __name__ == "__main__":
    main()
