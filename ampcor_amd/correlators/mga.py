"""The MGA correlator component, without pyre.

Restates pkg/ampcor/correlators/MGA.py:20-160: its traits (chip, padding,
refineMargin, refineFactor, zoomFactor and the coarse map) and ``estimate``,
which plans the pairs, opens the rasters and asks the strategy for the
offsets.  ``estimate`` returns the offsets and products (the reference drops
them and returns 0, MGA.py:91-98), and ``offset_field`` lays them out on the
tile grid with NaN for invalid pairs.
"""

import numpy as np

from .hip import HIP
from .plan import Grid, Plan


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
        ref = reference.open().raster
        tgt = target.open().raster
        offsets = worker.adjust(manager=self, rasters=(ref, tgt), plan=plan, channel=channel)
        self.plan = plan
        self.products = worker.products()
        return offsets

    def offset_field(self):
        """Dense (grid..., 4) float32 raster: dLine, dSample, coarse peak, SNR; NaN where invalid."""
        valid = self.plan.valid
        field = np.full((valid.size, 4), np.nan, dtype=np.float32)
        p = self.products
        field[valid, 0:2] = p["offsets"]
        field[valid, 2] = p["coarse_peak"]
        field[valid, 3] = p["snr"]
        return field.reshape(tuple(self.plan.tile) + (4,))
