"""Plan geometry (pkg/ampcor/correlators/UniformGrid.py:34-54, Plan.py:143-185,
pkg/ampcor/dom/SLC.py:55-76) restated without pyre."""

import numpy as np
import pytest

from ampcor_amd import SLC, Constant, Field, Grid, Plan, UniformGrid


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


def grid_points(shape, bounds):
    return UniformGrid(shape=shape).points(bounds=bounds)


def test_field_of_a_constant_is_the_constant_functor():
    pts = grid_points((4, 5), (400, 500))
    field = Field(points=pts, offsets=np.tile([3.4, -7.6], (len(pts), 1)))
    dom = grid_points((7, 3), (400, 500))
    assert tuple(field.codomain(dom)) == tuple(Constant(shift=(3, -8)).codomain(dom))


def test_field_interpolates_a_linear_field_exactly():
    # bilinear interpolation reproduces any field linear in (line, sample)
    pts = np.array(grid_points((5, 6), (500, 600)), dtype=np.float64)
    f = lambda p: np.stack([0.05 * p[..., 0] - 0.01 * p[..., 1] + 2.0,
                            -0.02 * p[..., 0] + 0.03 * p[..., 1] - 5.0], axis=-1)
    field = Field(points=pts, offsets=f(pts))
    q = np.array([(60, 77), (123, 401), (449, 549), (250, 300)], dtype=np.float64)
    for p in q:
        assert np.allclose(field.at(p), f(p), atol=1e-12)
    img = np.array(tuple(field.codomain([tuple(int(v) for v in p) for p in q])))
    assert np.array_equal(img, q.astype(np.int64) + np.floor(f(q) + 0.5).astype(np.int64))


def test_field_rounds_half_up_and_holds_its_edges():
    pts = grid_points((2, 2), (200, 200))          # nodes at 50 and 150
    off = np.array([[0.5, -0.5], [0.5, -0.5], [2.5, 1.5], [2.5, 1.5]])
    field = Field(points=pts, offsets=off)
    assert next(field.codomain([(0, 0)])) == (1, 0)            # held at the first row: (0.5, -0.5)
    assert next(field.codomain([(199, 10)])) == (199 + 3, 12)  # held at the last row: (2.5, 1.5)
    assert next(field.codomain([(100, 10)])) == (100 + 2, 11)  # midway: (1.5, 0.5)


def test_field_fills_holes_from_the_nearest_valid_node():
    pts = grid_points((3, 3), (300, 300))
    off = np.full((9, 2), np.nan)
    off[0] = (1.0, 2.0)
    off[8] = (-4.0, 6.0)
    field = Field(points=pts, offsets=off)
    # node (0, 1) is nearest to node (0, 0); node (2, 1) to node (2, 2)
    assert np.array_equal(field.grid[0, 1], [1.0, 2.0])
    assert np.array_equal(field.grid[2, 1], [-4.0, 6.0])
    # the centre is equidistant: the lowest row-major valid node wins
    assert np.array_equal(field.grid[1, 1], [1.0, 2.0])
    empty = Field(points=pts, offsets=np.full((9, 2), np.nan), fallback=(5, -1))
    assert next(empty.codomain([(10, 10)])) == (15, 9)


def test_field_rejects_scattered_points():
    with pytest.raises(ValueError):
        Field(points=[(0, 0), (1, 5), (7, 3)], offsets=np.zeros((3, 2)))


def test_plan_records_each_tiles_initial_shift():
    ref, tgt = SLC(shape=(400, 400)), SLC(shape=(400, 400))
    pts = grid_points((2, 2), (400, 400))
    field = Field(points=pts, offsets=[[1, 2], [3, 4], [5, 6], [7, 8]])
    coarse = Grid(domain=UniformGrid(shape=(2, 2)), functor=field)
    plan = Plan(correlator=Corr(32, 16), regmap=coarse.map(reference=ref), rasters=(ref, tgt))
    assert plan.shift.tolist() == [[1, 2], [3, 4], [5, 6], [7, 8]]
    r, t = plan.corners()
    assert np.array_equal(t - r, plan.shift - 16)


def test_next_pass_field_from_an_estimate_without_a_device():
    # MGA.offset_field / next_pass on products set by hand (no strategy runs)
    from ampcor_amd import MGA

    ref, tgt = SLC(shape=(400, 400)), SLC(shape=(400, 400))
    mga = MGA(chip=(32, 32), padding=(16, 16),
              coarse=Grid(domain=UniformGrid(shape=(2, 2)), functor=Constant(shift=(5, -3))))
    mga.plan = mga.makePlan(regmap=mga.coarse.map(reference=ref), rasters=(ref, tgt))
    n = len(mga.plan)
    assert n == 4
    mga.products = {"offsets": np.array([[0.25, -0.5], [0.5, 0.25], [-0.25, 0.0], [1.0, 1.0]],
                                        np.float32),
                    "coarse_peak": np.full(n, 0.9, np.float32), "snr": np.full(n, 30.0, np.float32)}
    rel = mga.offset_field()
    tot = mga.offset_field(total=True)
    assert rel.shape == (2, 2, 4)
    assert np.allclose(tot[..., 0:2] - rel[..., 0:2], [5, -3])
    coarse = mga.next_pass(shape=(2, 2))
    offmap = coarse.map(reference=ref)
    # the next pass's guesses on the same grid: the rounded total offsets
    got = np.array(offmap.codomain) - np.array(offmap.domain)
    want = np.floor(tot[..., 0:2].reshape(-1, 2) + 0.5).astype(int)
    assert np.array_equal(got, want)


def test_offset_map_is_indexed_row_major():
    # pkg/ampcor/dom/OffsetMap.py:14-95: map[index] is the (reference, target) pair at a grid index
    from ampcor_amd.correlators import OffsetMap

    m = Grid(domain=UniformGrid(shape=(2, 3)), functor=Constant(shift=(1, -1))).map(
        reference=SLC(shape=(100, 90)))
    assert m.shape == (2, 3) and m.size == 6 and m.layout == (1, 0)
    assert m[0, 0] == ((25, 15), (26, 14))
    assert m[1, 2] == ((75, 75), (76, 74))
    m[1, 2] = ((75, 75), (80, 70))
    assert m.codomain[5] == (80, 70)
    with pytest.raises(IndexError):
        m[2, 0]
    with pytest.raises(IndexError):
        m[0]
    assert OffsetMap(shape=(4, 5)).size == 20


def test_registration_map_of_an_estimate_without_a_device():
    from ampcor_amd import MGA

    ref, tgt = SLC(shape=(400, 400)), SLC(shape=(400, 400))
    mga = MGA(chip=(32, 32), padding=(16, 16),
              coarse=Grid(domain=UniformGrid(shape=(2, 2)), functor=Constant(shift=(5, -3))))
    mga.plan = mga.makePlan(regmap=mga.coarse.map(reference=ref), rasters=(ref, tgt))
    mga.products = {"offsets": np.array([[0.25, -0.5], [0.5, 0.25], [-0.25, 0.0], [1.0, 1.0]],
                                        np.float32),
                    "coarse_peak": np.full(4, 0.9, np.float32), "snr": np.full(4, 30.0, np.float32)}
    m = mga.offset_map()
    assert m.shape == (2, 2)
    r, t = m[0, 1]
    # pair (0, 1): offsets (0.5, 0.25) on the guess (5, -3)
    assert r == (100, 300) and t == (100 + 5.5, 300 - 3 + 0.25)


def test_plan_indexing_and_footprint():
    # Plan.py:48-59 (footprint) and :129-141 (grid indexing)
    plan = plan_for((512, 512), (8, 8), 32, 16, shift=(0, 0))
    r, t = plan[0, 0]
    assert r.begin == (16, 16) and t.begin == (0, 0)
    assert plan.footprint == (64 * 32 * 32 * 8, 64 * 64 * 64 * 8)
    edge = plan_for((100, 100), (2, 2), 32, 16)
    assert edge[0, 0] == (None, None)   # the window would start before the raster
    with pytest.raises(IndexError):
        plan[8, 0]


def test_slc_and_slice_pixels_from_host_data():
    z = (np.arange(12) + 1j * np.arange(12)[::-1]).astype(np.complex64).reshape(3, 4)
    slc = SLC(data=z)
    assert slc[1, 2] == complex(z[1, 2])
    assert slc.slice((1, 1), (2, 2))[1, 0] == complex(z[2, 1])
    with pytest.raises(IndexError):
        slc[3, 0]
