"""The MGA component on the GPU: estimate through the HIP strategy, the offset
raster on the tile grid, and an iterative second pass whose coarse map follows
the first pass's offsets (runs on the GPU).

Bar: the component returns exactly what the worker path returns on the same
pairs; the second pass's offsets are within 1e-3 px of the FP64 oracle on the
pairs that pass planned, and its total displacement recovers the planted
shift, which lies beyond the second pass's own search margin.
"""

import numpy as np
import pytest

from ampcor_amd import MGA, SLC, Grid, UniformGrid, synth
from oracle import pyoracle
from tests.gpu_helpers import run_hip

pytestmark = pytest.mark.gpu


def test_estimate_equals_the_worker_path_and_fills_the_tile_grid(lib):
    s = synth.shifted_scene(seed=3)
    print(f"MGA estimate: {s.pairs} pairs", flush=True)
    mga = MGA(chip=(s.chip, s.chip), padding=(s.margin, s.margin),
              refineMargin=s.refine_margin, refineFactor=s.refine_factor,
              zoomFactor=s.zoom_factor, coarse=Grid(domain=UniformGrid(shape=s.grid)))
    ref, tgt = SLC(data=s.reference), SLC(data=s.target)
    offsets = np.array(mga.estimate(ref, tgt), np.float32)
    # the opened rasters answer pixel reads from the device copy (SLC.py:102-110)
    assert ref.raster is not None
    assert ref[5, 7] == complex(s.reference[5, 7]) and tgt[300, 2] == complex(s.target[300, 2])
    assert mga.plan[0, 0][0][3, 4] == complex(s.reference[s.ref_begin[0][0] + 3,
                                                          s.ref_begin[0][1] + 4])
    direct = run_hip(lib, s.reference, s.target, s.ref_begin, s.tgt_begin, s.chip, s.window,
                     s.refine_factor, s.refine_margin, s.zoom_factor)
    assert np.array_equal(offsets.view(np.uint32), direct["offsets"].view(np.uint32))
    field = mga.offset_field()
    assert field.shape == tuple(s.grid) + (4,)
    flat = field.reshape(-1, 4)
    valid = mga.plan.valid
    assert np.array_equal(flat[valid, 0:2], direct["offsets"])
    assert np.array_equal(flat[valid, 2], direct["coarse_peak"])
    assert np.array_equal(flat[valid, 3], direct["snr"])
    assert np.isnan(flat[~valid]).all()
    # no initial shift: the total field is the offsets themselves
    assert np.array_equal(mga.offset_field(total=True), field)


def test_second_pass_follows_the_first_pass_field(lib):
    # a shift beyond the second pass's margin of 6 px: only the field finds it
    shift = (13.4, -11.3)
    s = synth.shifted_scene(seed=9, size=512, grid=(3, 3), chip=32, margin=24,
                            refine_margin=6, zoom_factor=4, shift=shift)
    ref, tgt = SLC(data=s.reference), SLC(data=s.target)
    first = MGA(chip=(32, 32), padding=(24, 24), refineMargin=6, refineFactor=2, zoomFactor=4,
                coarse=Grid(domain=UniformGrid(shape=(3, 3))))
    first.estimate(ref, tgt)
    assert np.all(np.abs(first.offset_field(total=True)[..., 0:2] - shift) < 0.05)

    second = MGA(chip=(32, 32), padding=(6, 6), refineMargin=6, refineFactor=2, zoomFactor=4,
                 coarse=first.next_pass(shape=(8, 8)))
    offsets = np.array(second.estimate(ref, tgt), np.float32).reshape(-1, 2)
    plan = second.plan
    print(f"second pass: {len(plan)} pairs of {plan.valid.size}", flush=True)
    assert len(plan) >= 36
    # every guess is the rounded shift (the field is constant up to its errors)
    assert np.array_equal(plan.shift, np.tile([13, -11], (plan.shift.shape[0], 1)))
    rb, tb = plan.corners()
    exact = pyoracle.adjust(s.reference, s.target, rb, tb, 32, 32 + 12, 2, 6, 4, form="fp64")
    np.testing.assert_allclose(offsets, exact["offsets"], rtol=0, atol=1e-3)
    total = second.offset_field(total=True)[..., 0:2].reshape(-1, 2)[plan.valid]
    assert np.all(np.abs(total - shift) < 0.05)
