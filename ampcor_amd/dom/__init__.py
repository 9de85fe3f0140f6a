"""Raster data model: SLC rasters and their slices."""

from .slc import SLC, Slice

__all__ = ["SLC", "Slice"]
