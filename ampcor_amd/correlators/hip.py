"""The HIP correlation strategy.

Drops in beside the reference's strategies (pkg/ampcor/correlators/CUDA.py:22-61,
Sequential.py:22-51): same ``adjust(manager, rasters, plan, channel)``
signature, same worker calls, same result -- a tuple of ``(dLine, dSample)``
per valid pair.  The device is an argument (the reference hardcodes device 4,
CUDA.py:73).
"""

import numpy as np

from .. import ext


class HIP:
    """The MI355X accelerated registration strategy."""

    def __init__(self, device=0, precision="fp32", batched=True):
        self.device = device
        self.precision = precision
        self.batched = batched
        self.worker = None
        self.lib = ext.require()

    def adjust(self, manager, rasters, plan, channel=None):
        """Correlate a pair of rasters given a plan; returns the offset tuple."""
        ref, tgt = rasters
        lib = self.lib
        points = len(plan)
        worker = lib.sequential(points, plan.chip, plan.window, manager.refineMargin,
                                manager.refineFactor, manager.zoomFactor,
                                precision=self.precision, device=self.device)
        if channel is not None:
            channel.log("loading tiles")
        if self.batched:
            # one call for all pairs; the same descriptors the per-pair calls record
            refBegin, tgtBegin = plan.corners()
            lib.addPairs(worker, ref, tgt, 0, refBegin.tobytes(), tgtBegin.tobytes())
        else:
            for idx, (r, t) in enumerate(plan.pairs):
                lib.addReference(worker, ref, idx, r.begin, r.end)
                lib.addTarget(worker, tgt, idx, t.begin, t.end)
        if channel is not None:
            channel.log("adjusting the offset map")
        offsets = lib.adjust(worker)
        self.worker = worker
        return offsets

    def products(self):
        """Offsets [N, 2], coarse peak, refined peak and SNR [N] of the last adjust."""
        if self.worker is None:
            raise RuntimeError("products requested before adjust")
        raw = self.lib.products(self.worker)
        out = {k: np.frombuffer(v, dtype=np.float32) for k, v in raw.items()}
        out["offsets"] = out["offsets"].reshape(-1, 2)
        return out
