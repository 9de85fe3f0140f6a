"""The oracle against the committed golden fixtures (tests/golden/make_golden.py)."""

import glob
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "upsample_*.npz"))))
def test_upsample_golden(path):
    f = np.load(path)
    np.testing.assert_allclose(po.upsample(f["x"], int(f["rf"])), f["y"], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "zoom_*.npz"))))
def test_zoom_golden(path):
    f = np.load(path)
    np.testing.assert_allclose(po.zoom_full(f["g"], int(f["zf"])), f["z"], rtol=1e-10, atol=1e-11)


def test_end_to_end_golden_shift():
    f = np.load(os.path.join(GOLDEN, "shift_e2e.npz"))
    C, W, rf, rm, zf = (int(v) for v in f["geometry"])
    out = po.adjust(f["reference"], f["target"], f["ref_begin"], f["tgt_begin"], C, W, rf, rm, zf)
    assert np.max(np.abs(out["offsets"] - f["shift"])) < 0.05


def test_end_to_end_golden_offgrid_shift():
    """A shift that is not a multiple of the zoomed cell (1/8 px): the sub-pixel
    step carries about 0.06 px of every offset, so the 0.03 px bar fails when it
    is left out or flipped."""
    f = np.load(os.path.join(GOLDEN, "shift_offgrid.npz"))
    C, W, rf, rm, zf = (int(v) for v in f["geometry"])
    out = po.adjust(f["reference"], f["target"], f["ref_begin"], f["tgt_begin"], C, W, rf, rm, zf)
    assert np.max(np.abs(out["offsets"] - f["shift"])) < 0.03
    m, z = (W - C) // 2, rf * zf
    grid_only = (out["window_loc"] - m) + out["zoom_loc"] / z
    assert np.min(np.abs(out["offsets"] - grid_only)) > 0.03
