"""Seeded synthetic scenes for tests and benchmarks.

Circular complex Gaussian speckle (unit variance), secondaries made by a
periodic Fourier shift or by resampling along a smooth, separable offset
field, complex Gaussian noise at a given SNR, and real lognormal texture.
Also the plan geometry of the standard workloads (a uniform grid of tile
centers, constant zero coarse shift), restated as arithmetic.
"""

import os
from dataclasses import dataclass

import numpy as np

from .correlators.plan import Constant, Grid, Plan, UniformGrid
from .dom.slc import SLC


def speckle(shape, rng):
    """Circular complex Gaussian, E|z|^2 = 1 (re, im ~ N(0, 1/2)), complex64."""
    z = rng.standard_normal(shape, dtype=np.float32) + 1j * rng.standard_normal(shape, dtype=np.float32)
    return (z * np.float32(np.sqrt(0.5))).astype(np.complex64)


def fourier_shift(x, dline, dsample):
    """Periodic band-limited shift: out(l, s) = x(l - dline, s - dsample)."""
    lines, samples = x.shape
    fy = np.fft.fftfreq(lines)[:, None]
    fx = np.fft.fftfreq(samples)[None, :]
    phase = np.exp(-2j * np.pi * (fy * dline + fx * dsample))
    out = np.fft.ifft2(np.fft.fft2(x.astype(np.complex128)) * phase)
    return out.astype(x.dtype if np.iscomplexobj(x) else np.complex64)


def add_noise(x, snr_db, rng):
    """Add circular Gaussian noise of power mean|x|^2 / 10^(snr/10)."""
    power = float(np.mean(np.abs(x) ** 2)) / (10.0 ** (snr_db / 10.0))
    if np.iscomplexobj(x):
        n = speckle(x.shape, rng) * np.float32(np.sqrt(power))
    else:
        n = rng.standard_normal(x.shape, dtype=np.float32) * np.float32(np.sqrt(power))
    return (x + n).astype(x.dtype)


def _sinc_weights(frac, taps):
    """Kaiser-windowed sinc weights for positions k - frac, k in [-taps/2 + 1, taps/2]."""
    k = np.arange(-taps // 2 + 1, taps // 2 + 1, dtype=np.float64)
    d = k[None, :] - frac[:, None]
    beta = 6.0
    window = np.i0(beta * np.sqrt(np.clip(1.0 - (d / (taps / 2)) ** 2, 0.0, None))) / np.i0(beta)
    w = np.sinc(d) * window
    return (w / w.sum(axis=1, keepdims=True)).astype(np.float32), k.astype(np.int64)


def resample_axis(x, shift, axis, taps=16, threads=None):
    """out[..., i, ...] = x interpolated at i - shift[i] along ``axis`` (edges clamped).

    Runs over bands of the other axis in a thread pool (NumPy releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    n = x.shape[axis]
    pos = np.arange(n, dtype=np.float64) - shift
    base = np.floor(pos).astype(np.int64)
    w, k = _sinc_weights(pos - base, taps)
    idx = [np.clip(base + k[t], 0, n - 1) for t in range(len(k))]
    out = np.empty_like(x)
    other = x.shape[1 - axis]
    threads = threads or min(16, len(os.sched_getaffinity(0)))
    step = max(64, -(-other // (4 * threads)))

    def band(lo):
        hi = min(lo + step, other)
        if axis == 0:
            xs, acc = x[:, lo:hi], np.zeros((n, hi - lo), dtype=x.dtype)
            for t in range(len(k)):
                acc += w[:, t, None] * xs[idx[t]]
            out[:, lo:hi] = acc
        else:
            xs, acc = x[lo:hi], np.zeros((hi - lo, n), dtype=x.dtype)
            for t in range(len(k)):
                acc += w[None, :, t] * xs[:, idx[t]]
            out[lo:hi] = acc

    with ThreadPoolExecutor(threads) as pool:
        list(pool.map(band, range(0, other, step)))
    return out


def lognormal_texture(shape, sigma, rng):
    """exp(sigma N(0,1)) smoothed by a 5x5 Gaussian (std 1 px), float32."""
    t = np.exp(sigma * rng.standard_normal(shape, dtype=np.float32)).astype(np.float32)
    g = np.exp(-0.5 * np.arange(-2, 3) ** 2)
    g = (g / g.sum()).astype(np.float32)
    t = sum(g[i] * np.roll(t, i - 2, axis=0) for i in range(5))
    t = sum(g[i] * np.roll(t, i - 2, axis=1) for i in range(5))
    return t.astype(np.float32)


@dataclass
class Scene:
    """A raster pair plus the plan of a workload."""

    name: str
    reference: np.ndarray
    target: np.ndarray
    chip: int
    margin: int
    refine_factor: int
    refine_margin: int
    zoom_factor: int
    grid: tuple
    ref_begin: np.ndarray = None
    tgt_begin: np.ndarray = None
    valid: np.ndarray = None
    truth: np.ndarray = None  # true (dLine, dSample) per valid pair, when known

    @property
    def window(self):
        return self.chip + 2 * self.margin

    @property
    def pairs(self):
        return len(self.ref_begin)

    def plan(self):
        """Restate Plan.assemble over a uniform grid with no coarse shift."""
        ref = SLC(shape=self.reference.shape)
        tgt = SLC(shape=self.target.shape)
        coarse = Grid(domain=UniformGrid(shape=self.grid), functor=Constant(shift=(0, 0)))

        class _Correlator:
            chip = (self.chip, self.chip)
            padding = (self.margin, self.margin)

        plan = Plan(correlator=_Correlator, regmap=coarse.map(reference=ref), rasters=(ref, tgt))
        self.ref_begin, self.tgt_begin = plan.corners()
        self.valid = plan.valid
        return plan


def planted(chip, margin, seed=0):
    """The planted-chip fixture of the reference's CUDA tests
    (tests/libampcor_cuda/correlate.cc:116-145): one pair per placement (i, j);
    pair (i, j) holds a real N(0,1) chip copied into an all-zero window at (i, j).
    Returns one raster holding all chips side by side and one holding all windows."""
    rng = np.random.default_rng(seed)
    P = 2 * margin + 1
    W = chip + 2 * margin
    chips = rng.standard_normal((chip, chip)).astype(np.float32)
    n = P * P
    ref = np.zeros((chip, n * chip), dtype=np.complex64)
    tgt = np.zeros((W, n * W), dtype=np.complex64)
    for i in range(P):
        for j in range(P):
            k = i * P + j
            ref[:, k * chip:(k + 1) * chip] = chips
            tgt[i:i + chip, k * W + j:k * W + j + chip] = chips
    ref_begin = np.array([(0, k * chip) for k in range(n)], dtype=np.int64)
    tgt_begin = np.array([(0, k * W) for k in range(n)], dtype=np.int64)
    return ref, tgt, ref_begin, tgt_begin


def shifted_scene(seed=1, size=512, grid=(8, 8), chip=32, margin=16, refine_factor=2,
                  refine_margin=8, zoom_factor=4, shift=(3.25, -1.75), snr_db=20.0):
    """Speckle reference; secondary = reference Fourier-shifted by ``shift`` plus noise."""
    rng = np.random.default_rng(seed)
    ref = speckle((size, size), rng)
    tgt = add_noise(fourier_shift(ref, *shift), snr_db, rng)
    s = Scene("shifted", ref, tgt, chip, margin, refine_factor, refine_margin, zoom_factor, grid)
    s.plan()
    s.truth = np.tile(np.array(shift, dtype=np.float64), (s.pairs, 1))
    return s


def field_scene(seed=2, lines=8192, samples=4096, grid=(120, 60), chip=64, margin=32,
                refine_factor=2, refine_margin=8, zoom_factor=8, range_ramp=5.0,
                azimuth_quadratic=2.0, snr_db=15.0, taps=16):
    """Speckle reference; secondary resampled by a smooth separable field:
    dSample = range_ramp * s / (samples - 1), dLine = azimuth_quadratic * (l / (lines - 1))^2,
    plus noise."""
    rng = np.random.default_rng(seed)
    ref = speckle((lines, samples), rng)
    ds = range_ramp * np.arange(samples) / (samples - 1)
    dl = azimuth_quadratic * (np.arange(lines) / (lines - 1)) ** 2
    tgt = resample_axis(resample_axis(ref, dl, axis=0, taps=taps), ds, axis=1, taps=taps)
    tgt = add_noise(tgt, snr_db, rng)
    s = Scene("field", ref, tgt, chip, margin, refine_factor, refine_margin, zoom_factor, grid)
    s.plan()
    centers = s.ref_begin + chip // 2
    s.truth = np.stack([dl[centers[:, 0]], ds[centers[:, 1]]], axis=1)
    return s


def amplitude_scene(seed=4, size=8192, grid=(64, 32), chip=128, margin=64, refine_factor=2,
                    refine_margin=8, zoom_factor=16, max_shift=12.0, sigma=0.5, noise=0.1, taps=8):
    """Real amplitude rasters: lognormal texture (sigma) smoothed by a 5 x 5
    Gaussian; the secondary is resampled by a random smooth separable field of
    up to max_shift px and multiplied by (1 + noise N(0, 1))."""
    rng = np.random.default_rng(seed)
    ref = lognormal_texture((size, size), sigma, rng)
    a, b = rng.uniform(0.5, 1.0, size=2) * max_shift * rng.choice([-1.0, 1.0], size=2)
    x = np.arange(size) / (size - 1)
    dl = a * np.sin(0.5 * np.pi * x)          # smooth, monotone, up to |a| px
    ds = b * x ** 2                           # smooth, up to |b| px
    tgt = resample_axis(resample_axis(ref, dl, axis=0, taps=taps), ds, axis=1, taps=taps)
    tgt = (tgt * (1.0 + noise * rng.standard_normal(tgt.shape, dtype=np.float32))).astype(np.float32)
    s = Scene("amplitude", ref, tgt, chip, margin, refine_factor, refine_margin, zoom_factor, grid)
    s.plan()
    centers = s.ref_begin + chip // 2
    s.truth = np.stack([dl[centers[:, 0]], ds[centers[:, 1]]], axis=1)
    return s
