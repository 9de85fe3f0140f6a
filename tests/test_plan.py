"""Plan geometry (pkg/ampcor/correlators/UniformGrid.py:34-54, Plan.py:143-185,
pkg/ampcor/dom/SLC.py:55-76) restated without pyre."""

import numpy as np
import pytest

from ampcor_amd import SLC, Constant, Grid, Plan, UniformGrid


class Corr:
    def __init__(self, chip, padding):
        self.chip = (chip, chip)
        self.padding = (padding, padding)


def plan_for(shape, grid, chip, padding, shift=(0, 0)):
    ref, tgt = SLC(shape=shape), SLC(shape=shape)
    coarse = Grid(domain=UniformGrid(shape=grid), functor=Constant(shift=shift))
    return Plan(correlator=Corr(chip, padding), regmap=coarse.map(reference=ref), rasters=(ref, tgt))


@pytest.mark.parametrize("shape,grid,chip,pad,valid", [
    ((512, 512), (8, 8), 32, 16, 64),          # 512^2 scene, 8x8 grid
    ((8192, 4096), (120, 60), 64, 32, 6844),   # scene-sized, 7200 candidates
    ((8192, 8192), (64, 32), 128, 64, 1984),   # large-chip amplitude case
])
def test_valid_pair_counts(shape, grid, chip, pad, valid):
    plan = plan_for(shape, grid, chip, pad)
    assert len(plan) == valid
    assert plan.valid.sum() == valid and plan.valid.size == grid[0] * grid[1]


def test_centers_and_corners():
    pts = UniformGrid(shape=(2, 3)).points(bounds=(100, 90))
    # tile = (50, 30), margin = (0, 0): centers m//2 + n t + t//2
    assert pts == ((25, 15), (25, 45), (25, 75), (75, 15), (75, 45), (75, 75))
    plan = plan_for((100, 90), (2, 3), 10, 5)
    ref, tgt = plan.corners()
    assert tuple(ref[0]) == (20, 10) and tuple(tgt[0]) == (15, 5)
    r, t = next(iter(plan.pairs))
    assert r.end == (30, 20) and t.end == (35, 25)


def test_slices_past_the_raster_are_invalid():
    s = SLC(shape=(10, 10))
    assert s.slice((0, 0), (10, 10)) is not None
    assert s.slice((-1, 0), (2, 2)) is None
    assert s.slice((5, 5), (6, 5)) is None


def test_shifted_coarse_map_moves_the_windows():
    plan = plan_for((512, 512), (4, 4), 32, 16, shift=(3, -2))
    ref, tgt = plan.corners()
    assert np.all(tgt - ref == np.array([3 - 16, -2 - 16]))
