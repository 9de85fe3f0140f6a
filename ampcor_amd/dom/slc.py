"""SLC rasters and slices, without pyre.

Restates pkg/ampcor/dom/SLC.py:40-89 and pkg/ampcor/dom/Slice.py:49-67 of the
reference: an SLC is a (lines, samples) complex64 raster; ``slice`` returns
``None`` for a slice that starts at a negative index or runs past the raster.
``open`` places the raster in device memory (the reference memory-maps it,
lib/libampcor/dom/SLC.icc:122-127) and leaves its handle in ``raster``.
"""

import functools
import operator

import numpy as np

from .. import ext


class Slice:
    """A rectangular region of a raster: ``begin`` and one-past-the-end ``end``."""

    def __init__(self, raster, begin, shape):
        self.raster = raster
        self.begin = tuple(begin)
        self.end = tuple(b + s for b, s in zip(begin, shape))
        self.shape = tuple(shape)
        self.size = functools.reduce(operator.mul, shape, 1)
        self.footprint = self.size * raster.pixelSize

    def __repr__(self):
        return f"Slice(begin={self.begin}, shape={self.shape})"

    def __getitem__(self, index):
        """The raster's pixel at ``index`` relative to the slice (Slice.py:70-77)."""
        return self.raster[tuple(b + i for b, i in zip(self.begin, index))]


class SLC:
    """A complex64 (or float32 amplitude) raster of shape (lines, samples)."""

    pixelSize = 8

    def __init__(self, shape=None, data=None, path=None, dtype="complex64", device=0):
        if data is not None:
            data = np.ascontiguousarray(data)
            shape = data.shape
            dtype = {np.dtype(np.complex64): "complex64",
                     np.dtype(np.float32): "float32"}[data.dtype]
        if shape is None:
            raise ValueError("an SLC needs a shape or data")
        self.shape = tuple(int(s) for s in shape)
        self.data = data
        self.path = path
        self.dtype = dtype
        self.device = device
        self.pixelSize = 8 if dtype == "complex64" else 4
        self.raster = None

    def size(self):
        lines, samples = self.shape
        return lines * samples * self.pixelSize

    def __getitem__(self, index):
        """The pixel at a (line, sample) index, as a Python complex (SLC.py:102-110).

        Read from the device copy once the raster is open (``slc_getitem`` of
        the extension, at the row-major offset), else from the host data."""
        line, sample = index
        lines, samples = self.shape
        if not (0 <= line < lines and 0 <= sample < samples):
            raise IndexError(f"index {index} outside the raster {self.shape}")
        if self.raster is not None:
            return ext.require().slc_getitem(self.raster, line * samples + sample)
        if self.data is None:
            raise ValueError("the raster has no data: open it or give it data")
        return complex(self.data[line, sample])

    def slice(self, begin, shape):
        """A slice, or ``None`` when it starts before or ends past the raster."""
        for b in begin:
            if b < 0:
                return None
        for b, s, l in zip(begin, shape, self.shape):
            if b + s > l:
                return None
        return Slice(raster=self, begin=begin, shape=shape)

    def open(self, mode="r"):
        """Place the raster in device memory; its handle lands in ``self.raster``."""
        lib = ext.require()
        lines, samples = self.shape
        if self.data is not None:
            self.raster = lib.raster(memoryview(self.data).cast("B"), lines, samples,
                                     dtype=self.dtype, device=self.device)
        else:
            self.raster = lib.slc_map(filename=str(self.path), samples=samples, lines=lines,
                                      device=self.device)
        return self
