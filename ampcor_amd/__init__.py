"""ampcor_amd -- an MI355X (gfx950) batched tile correlator for ampcor.

The hot path of ampcor's offset tracking -- normalized cross-correlation of
reference chips against target search windows, peak finding, band-limited
refinement, zoom and sub-pixel offsets -- runs as hand-written HIP kernels in
``libampcor_hip.so`` behind a C ABI (``include/ampcor_hip.h``).  The CPython
module ``ampcor_hip`` mirrors the reference's ``ampcor_cuda`` extension, and
``correlators.HIP`` drops in beside the reference's ``CUDA`` and
``Sequential`` strategies.
"""

from . import ext
from .correlators import HIP, MGA, Constant, Field, Grid, Plan, UniformGrid
from .dom import SLC, Slice

__all__ = ["ext", "HIP", "MGA", "Constant", "Field", "Grid", "Plan", "UniformGrid", "SLC", "Slice"]
