"""The zoom stage (k_zoom) on given refined surfaces, against the full-grid zoom.

The reference zooms the whole refined surface and takes the maximum of every
zoomed cell (Sequential.icc:122-124).  The HIP stage evaluates a window
around the surface's maximum exactly and screens the rest of the grid on FP16
matrix cores, evaluating exactly whatever the screen cannot rule out.  Its
zoomed location, peak and sub-pixel offset must equal the oracle's full-grid
search (oracle_zoom_search) bitwise, on:
- the committed golden surfaces (tests/golden/zoom_*.npz, NumPy FFT zoom);
- random surfaces, whose highest zoomed cell is often on another lobe than
  the highest sample;
- refined surfaces of known-shift scenes.
"""

import glob
import os

import numpy as np
import pytest

from ampcor_amd import synth
from oracle import pyoracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def zoom_plan(lib, pairs, R, zf):
    """A plan whose refined surface is R x R (rf 1, rm (R - 1) / 2) with zoom zf."""
    rm = (R - 1) // 2
    chip = 16
    return lib.sequential(pairs, (chip, chip), (chip + 2 * rm, chip + 2 * rm), rm, 1, zf)


def run_zoom(lib, surfaces, zf):
    surfaces = np.ascontiguousarray(surfaces, dtype=np.float32)
    n, R, _ = surfaces.shape
    worker = zoom_plan(lib, n, R, zf)
    lib.stage_zoom(worker, surfaces.tobytes())
    prods = {k: np.frombuffer(v, np.float32) for k, v in lib.products(worker).items()}
    loc = np.frombuffer(lib.stage_locations(worker)["zoomed"], np.int32).reshape(n, 2)
    return loc, prods["refined_peak"], prods["offsets"].reshape(n, 2), lib.zoom_stats(worker)


def oracle_zoom(surfaces, zf):
    R = surfaces.shape[1]
    rm = (R - 1) // 2
    loc, peak, off = [], [], []
    for g in surfaces:
        (p, q), pk, (dp, dq) = pyoracle.zoom_search(g.astype(np.float64), zf)
        loc.append((p, q))
        peak.append(np.float32(pk))
        # the stage's offsets for a window at (0, 0): (p + dp) / zf - rm (rf = 1)
        off.append((np.float32((0 - rm) + (p + dp) / zf), np.float32((0 - rm) + (q + dq) / zf)))
    return np.array(loc, np.int32), np.array(peak, np.float32), np.array(off, np.float32)


def check(lib, surfaces, zf):
    loc, peak, off, stats = run_zoom(lib, surfaces, zf)
    eloc, epeak, eoff = oracle_zoom(surfaces, zf)
    print(f"zoom R={surfaces.shape[1]} zf={zf}: {len(surfaces)} surfaces, screen sent "
          f"{stats['cells']} cells of {stats['pairs']} surfaces to exact evaluation", flush=True)
    assert np.array_equal(loc, eloc)
    assert np.array_equal(peak.view(np.uint32), epeak.view(np.uint32))
    assert np.array_equal(off.view(np.uint32), eoff.view(np.uint32))
    return loc, stats


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "zoom_*.npz"))))
def test_zoom_golden(lib, path):
    f = np.load(path)
    g, zf, z = f["g"], int(f["zf"]), f["z"]
    loc, _ = check(lib, g[None].astype(np.float32), zf)
    # the full zoom of the same (float32) surface has its maximum on the same cell
    full = pyoracle.zoom_full(g.astype(np.float32).astype(np.float64), zf)
    np.testing.assert_allclose(full, z, atol=1e-6)
    assert tuple(loc[0]) == np.unravel_index(np.argmax(full), full.shape)


@pytest.mark.parametrize("R,zf", [(33, 8), (33, 16), (17, 4)])
def test_zoom_random_surfaces(lib, R, zf):
    rng = np.random.default_rng(R * 100 + zf)
    g = rng.uniform(-1.0, 1.0, size=(64, R, R)).astype(np.float32)
    # a few smooth surfaces too, whose peak sits between samples
    y, x = np.mgrid[0:R, 0:R]
    for k in range(16):
        c = rng.uniform(R / 4, 3 * R / 4, size=2)
        w = rng.uniform(1.0, 4.0)
        g[k] = np.exp(-((y - c[0]) ** 2 + (x - c[1]) ** 2) / (2 * w * w)) + 0.01 * g[k]
    loc, stats = check(lib, g, zf)
    # most random surfaces have their zoomed maximum more than a sample away
    # from their largest sample: the full-grid search is exercised
    a = np.array([np.unravel_index(np.argmax(s), s.shape) for s in g])
    far = np.any(np.abs(loc - zf * a) > zf, axis=1).sum()
    print(f"zoom R={R} zf={zf}: {far} maxima more than one sample from the largest sample",
          flush=True)
    assert far > 0 and stats["cells"] > 0


def test_zoom_refined_surfaces_of_shifted_scenes(lib):
    surfaces = []
    for seed, shift in [(31, (3.31, -1.69)), (32, (-2.47, 0.83)), (33, (0.52, 2.95))]:
        s = synth.shifted_scene(seed=seed, size=256, grid=(4, 4), shift=shift)
        out = pyoracle.adjust(s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                              s.refine_factor, s.refine_margin, s.zoom_factor, surfaces=True)
        surfaces.append(out["gamma_refined"])
    check(lib, np.concatenate(surfaces), 4)


def test_zoom_listed_cells_and_full_search(lib):
    """Both ways the screen hands cells on: broad smooth peaks at zoom 16 leave
    a few cells per surface near the window for exact evaluation (listed), and
    surfaces of independent noise leave more than the list holds (the full
    search); either way the result is the full-grid search's."""
    R, zf = 33, 16
    rng = np.random.default_rng(77)
    y, x = np.mgrid[0:R, 0:R]
    broad = np.empty((24, R, R), np.float32)
    for k in range(len(broad)):
        c = rng.uniform(R / 3, 2 * R / 3, size=2)
        w = rng.uniform(3.0, 6.0)
        broad[k] = np.exp(-((y - c[0]) ** 2 + (x - c[1]) ** 2) / (2 * w * w))
        broad[k] += 1e-3 * rng.standard_normal((R, R)).astype(np.float32)
    _, stats = check(lib, broad, zf)
    print(f"broad peaks: {stats}", flush=True)
    # (a searched surface counts among the pairs with exactly evaluated cells too)
    assert stats["pairs"] - stats["searched"] >= len(broad) // 2
    noise = rng.uniform(-1.0, 1.0, size=(16, R, R)).astype(np.float32)
    _, stats = check(lib, noise, zf)
    print(f"noise: {stats}", flush=True)
    assert stats["searched"] > 0
