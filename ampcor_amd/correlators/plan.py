"""Plan geometry, without pyre.

Restates the caller side of the hot path from the reference:
  UniformGrid.points   pkg/ampcor/correlators/UniformGrid.py:34-54
  Constant.codomain    pkg/ampcor/correlators/Constant.py:32-43
  Grid.map             pkg/ampcor/correlators/Grid.py:35-50
  Plan / assemble      pkg/ampcor/correlators/Plan.py:15-185
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


class OffsetMap:
    """A coarse offset map: reference points (domain) and their target images."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.tile = self.shape
        self.domain = ()
        self.codomain = ()


class Grid:
    """The initial guess: a domain of reference points and a functor."""

    def __init__(self, domain=None, functor=None):
        self.domain = domain if domain is not None else UniformGrid()
        self.functor = functor if functor is not None else Constant()

    def map(self, reference):
        offmap = OffsetMap(shape=self.domain.shape)
        offmap.domain = tuple(self.domain.points(bounds=reference.shape))
        offmap.codomain = tuple(self.functor.codomain(domain=offmap.domain))
        return offmap


class Plan:
    """The tile pairs to correlate to refine a coarse offset map."""

    def __init__(self, correlator, regmap, rasters):
        self.tile = regmap.tile
        self.chip = tuple(correlator.chip)
        self.padding = tuple(correlator.padding)
        self.window = tuple(c + 2 * p for c, p in zip(self.chip, self.padding))
        self.reference, self.target = self.assemble(regmap=regmap, rasters=rasters)

    @property
    def pairs(self):
        for ref, tgt in zip(self.reference, self.target):
            if ref and tgt:
                yield ref, tgt

    def __len__(self):
        return len(tuple(filter(None, self.reference)))

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
