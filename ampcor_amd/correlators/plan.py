"""Plan geometry, without pyre.

Restates the caller side of the hot path from the reference:
  UniformGrid.points   pkg/ampcor/correlators/UniformGrid.py:34-54
  Constant.codomain    pkg/ampcor/correlators/Constant.py:32-43
  Grid.map             pkg/ampcor/correlators/Grid.py:35-50
  Plan / assemble      pkg/ampcor/correlators/Plan.py:15-185
  Functor protocol     pkg/ampcor/correlators/Functor.py:15-38 (``Field`` is a
                       second implementation beside ``Constant``)
Pair ``idx`` is the position of a valid pair in ``plan.pairs`` (invalid pairs,
whose chip or window leaves a raster, are dropped), exactly as the reference
numbers them when it feeds its workers (CUDA.py:50-54).
"""

import itertools

import numpy as np


class UniformGrid:
    """Tile centers on a uniform grid of ``shape`` over a raster of ``bounds``."""

    def __init__(self, shape=(1, 1)):
        self.shape = tuple(shape)

    def points(self, bounds):
        tile = tuple(b // s for b, s in zip(bounds, self.shape))
        margin = tuple(b % s for b, s in zip(bounds, self.shape))
        ticks = tuple(
            tuple(m // 2 + n * t + t // 2 for n in range(g))
            for g, m, t in zip(self.shape, margin, tile))
        return tuple(itertools.product(*ticks))


class Constant:
    """Maps reference points to target points by a constant integer shift."""

    def __init__(self, shift=(0, 0)):
        self.shift = tuple(shift)

    def codomain(self, domain):
        for point in domain:
            yield tuple(p + s for p, s in zip(point, self.shift))


class Field:
    """Maps reference points through a dense offset field: the coarse map of an
    iterative pass, built from the offsets of the pass before it.

    ``points`` are the (line, sample) reference points of the earlier pass, on
    a rectilinear grid (what ``UniformGrid.points`` yields), and ``offsets``
    their total displacements (dLine, dSample) into the target, NaN where the
    earlier pass had no valid pair.  The image of a point is the point plus
    the field interpolated bilinearly at it, each coordinate rounded half up
    to an integer (window corners are whole pixels).  Outside the grid the
    field is held at its edge value.  A NaN node takes the value of the
    nearest valid node (grid-index distance, ties to the lowest row-major
    index); a field with no valid node maps every point by ``fallback``.
    """

    def __init__(self, points, offsets, fallback=(0, 0)):
        pts = np.asarray(points, dtype=np.float64).reshape(-1, 2)
        off = np.asarray(offsets, dtype=np.float64).reshape(-1, 2)
        if pts.shape != off.shape:
            raise ValueError("points and offsets must have the same number of (line, sample) rows")
        self.lines = np.unique(pts[:, 0])
        self.samples = np.unique(pts[:, 1])
        if self.lines.size * self.samples.size != pts.shape[0]:
            raise ValueError("the field's points must form a full rectilinear grid")
        order = np.lexsort((pts[:, 1], pts[:, 0]))     # row-major over (line, sample)
        grid = off[order].reshape(self.lines.size, self.samples.size, 2)
        self.fallback = tuple(int(f) for f in fallback)
        self.grid = self._fill(grid)

    @staticmethod
    def _fill(grid):
        bad = np.isnan(grid).any(axis=2)
        if not bad.any() or bad.all():
            return None if bad.all() else grid
        good = np.argwhere(~bad)
        out = grid.copy()
        for i, j in np.argwhere(bad):
            d = (good[:, 0] - i) ** 2 + (good[:, 1] - j) ** 2
            gi, gj = good[int(np.argmin(d))]
            out[i, j] = grid[gi, gj]
        return out

    def _axis(self, ticks, x):
        """Cell index and weight of x on one axis, held at the ends."""
        if ticks.size == 1:
            return 0, 0, 0.0
        x = min(max(x, ticks[0]), ticks[-1])
        k = int(np.searchsorted(ticks, x, side="right")) - 1
        k = min(k, ticks.size - 2)
        return k, k + 1, (x - ticks[k]) / (ticks[k + 1] - ticks[k])

    def at(self, point):
        """The field (dLine, dSample) interpolated at a reference point."""
        if self.grid is None:
            return np.array(self.fallback, dtype=np.float64)
        i0, i1, a = self._axis(self.lines, float(point[0]))
        j0, j1, b = self._axis(self.samples, float(point[1]))
        g = self.grid
        return ((1 - a) * ((1 - b) * g[i0, j0] + b * g[i0, j1])
                + a * ((1 - b) * g[i1, j0] + b * g[i1, j1]))

    def codomain(self, domain):
        for point in domain:
            d = self.at(point)
            yield tuple(int(p) + int(np.floor(v + 0.5)) for p, v in zip(point, d))


class OffsetMap:
    """A logically Cartesian map from reference points (domain) to their target
    images (codomain), as pkg/ampcor/dom/OffsetMap.py:14-95: ``map[index]`` is
    the pair of points at a grid index, stored row-major (the last index
    varies fastest)."""

    def __init__(self, shape):
        self.shape = tuple(int(s) for s in shape)
        self.tile = self.shape
        self.layout = tuple(reversed(range(len(self.shape))))
        n = self.size
        self.domain = [None] * n
        self.codomain = [None] * n

    @property
    def size(self):
        n = 1
        for s in self.shape:
            n *= s
        return n

    def offset(self, index):
        """Row-major offset of a grid index; IndexError outside the grid."""
        index = tuple(index) if isinstance(index, (tuple, list)) else (index,)
        if len(index) != len(self.shape):
            raise IndexError(f"index {index} has the wrong rank for shape {self.shape}")
        off = 0
        for i, s in zip(index, self.shape):
            if not 0 <= i < s:
                raise IndexError(f"index {index} outside the grid {self.shape}")
            off = off * s + i
        return off

    def __getitem__(self, index):
        off = self.offset(index)
        return self.domain[off], self.codomain[off]

    def __setitem__(self, index, points):
        off = self.offset(index)
        ref, tgt = points
        self.domain[off] = ref
        self.codomain[off] = tgt


class Grid:
    """The initial guess: a domain of reference points and a functor."""

    def __init__(self, domain=None, functor=None):
        self.domain = domain if domain is not None else UniformGrid()
        self.functor = functor if functor is not None else Constant()

    def map(self, reference):
        offmap = OffsetMap(shape=self.domain.shape)
        offmap.domain = list(self.domain.points(bounds=reference.shape))
        offmap.codomain = list(self.functor.codomain(domain=offmap.domain))
        return offmap


class Plan:
    """The tile pairs to correlate to refine a coarse offset map."""

    def __init__(self, correlator, regmap, rasters):
        self.tile = regmap.tile
        self.chip = tuple(correlator.chip)
        self.padding = tuple(correlator.padding)
        self.window = tuple(c + 2 * p for c, p in zip(self.chip, self.padding))
        self.domain = tuple(regmap.domain)
        # the integer shift of each tile's initial guess: target point - reference point
        self.shift = np.array([tuple(t - r for r, t in zip(rp, tp))
                               for rp, tp in zip(regmap.domain, regmap.codomain)],
                              dtype=np.int64).reshape(-1, 2)
        self.reference, self.target = self.assemble(regmap=regmap, rasters=rasters)

    @property
    def pairs(self):
        for ref, tgt in zip(self.reference, self.target):
            if ref and tgt:
                yield ref, tgt

    def __len__(self):
        return len(tuple(filter(None, self.reference)))

    def __getitem__(self, index):
        """The (reference, target) slices at a tile-grid index (None, None when invalid)."""
        off = OffsetMap(shape=self.tile).offset(index)
        return self.reference[off], self.target[off]

    @property
    def footprint(self):
        """Bytes of the valid reference and target tiles (Plan.py:48-59)."""
        ref = sum(t.footprint for t in filter(None, self.reference))
        tgt = sum(t.footprint for t in filter(None, self.target))
        return ref, tgt

    @property
    def valid(self):
        """Boolean mask over the tile grid (row-major) of the valid pairs."""
        return np.array([bool(r and t) for r, t in zip(self.reference, self.target)])

    def corners(self):
        """int64 (line, sample) upper-left corners of the valid chips and windows."""
        ref = np.array([r.begin for r, _ in self.pairs], dtype=np.int64).reshape(-1, 2)
        tgt = np.array([t.begin for _, t in self.pairs], dtype=np.int64).reshape(-1, 2)
        return ref, tgt

    def assemble(self, rasters, regmap):
        reference, target = rasters
        chip, padding = self.chip, self.padding
        referenceTiles, targetTiles = [], []
        for ref, tgt in zip(regmap.domain, regmap.codomain):
            begin = tuple(r - c // 2 for r, c in zip(ref, chip))
            refSlice = reference.slice(begin=begin, shape=chip)
            begin = tuple(t - c // 2 - p for t, c, p in zip(tgt, chip, padding))
            shape = tuple(c + 2 * p for c, p in zip(chip, padding))
            tgtSlice = target.slice(begin=begin, shape=shape)
            if refSlice and tgtSlice:
                referenceTiles.append(refSlice)
                targetTiles.append(tgtSlice)
            else:
                referenceTiles.append(None)
                targetTiles.append(None)
        return referenceTiles, targetTiles
