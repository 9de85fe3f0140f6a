"""The MGA correlator component, without pyre.

Restates pkg/ampcor/correlators/MGA.py:20-160: its traits (chip, padding,
refineMargin, refineFactor, zoomFactor and the coarse map) and ``estimate``,
which plans the pairs, opens the rasters and asks the strategy for the
offsets.  ``estimate`` returns the offsets and products (the reference drops
them and returns 0, MGA.py:91-98), and ``offset_field`` lays them out on the
tile grid with NaN for invalid pairs.  ``next_pass`` turns the result into the
coarse map of a further, finer pass (a ``Field`` functor): the offsets the
strategy returns are relative to the initial guess, as in the reference, and
the field adds the guess back.
"""

import numpy as np

from .hip import HIP
from .plan import Field, Grid, OffsetMap, Plan, UniformGrid


class MGA:
    def __init__(self, chip=(128, 128), padding=(32, 32), refineMargin=8, refineFactor=2,
                 zoomFactor=8, coarse=None, device=0, precision="fp32"):
        self.chip = tuple(chip)
        self.padding = tuple(padding)
        self.refineMargin = refineMargin
        self.refineFactor = refineFactor
        self.zoomFactor = zoomFactor
        self.coarse = coarse if coarse is not None else Grid()
        self.device = device
        self.precision = precision

    def makeWorker(self):
        return HIP(device=self.device, precision=self.precision)

    def makePlan(self, regmap, rasters):
        return Plan(correlator=self, regmap=regmap, rasters=rasters)

    def estimate(self, reference, target, channel=None):
        worker = self.makeWorker()
        coarse = self.coarse.map(reference=reference)
        plan = self.makePlan(regmap=coarse, rasters=(reference, target))
        # rasters stay resident across passes: each is uploaded on its first open
        ref = (reference.raster if reference.raster is not None else reference.open().raster)
        tgt = (target.raster if target.raster is not None else target.open().raster)
        offsets = worker.adjust(manager=self, rasters=(ref, tgt), plan=plan, channel=channel)
        self.plan = plan
        self.products = worker.products()
        return offsets

    def offset_field(self, total=False):
        """Dense (grid..., 4) float32 raster: dLine, dSample, coarse peak, SNR; NaN where invalid.

        With ``total`` the offsets include the initial guess of each tile (the
        coarse map's integer shift), i.e. the full displacement into the target."""
        valid = self.plan.valid
        field = np.full((valid.size, 4), np.nan, dtype=np.float32)
        p = self.products
        off = p["offsets"].astype(np.float64)
        if total:
            off = off + self.plan.shift[valid]
        field[valid, 0:2] = off
        field[valid, 2] = p["coarse_peak"]
        field[valid, 3] = p["snr"]
        return field.reshape(tuple(self.plan.tile) + (4,))

    def offset_map(self):
        """The registration map this estimate found: an ``OffsetMap`` on the tile
        grid whose codomain is each reference point plus its total offset (float
        target positions; NaN for invalid pairs)."""
        total = self.offset_field(total=True)[..., 0:2].astype(np.float64).reshape(-1, 2)
        offmap = OffsetMap(shape=self.plan.tile)
        offmap.domain = list(self.plan.domain)
        offmap.codomain = [tuple(float(p) + d for p, d in zip(point, off))
                           for point, off in zip(self.plan.domain, total)]
        return offmap

    def next_pass(self, shape, fallback=(0, 0)):
        """The coarse map of a further pass: a ``shape`` grid of tiles whose
        initial guesses follow this pass's total offsets."""
        total = self.offset_field(total=True)[..., 0:2].astype(np.float64).reshape(-1, 2)
        functor = Field(points=self.plan.domain, offsets=total, fallback=fallback)
        return Grid(domain=UniformGrid(shape=shape), functor=functor)
