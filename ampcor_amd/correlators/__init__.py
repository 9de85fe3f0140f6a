"""Correlation strategies and the plan that feeds them."""

from .hip import HIP
from .mga import MGA
from .plan import Constant, Grid, OffsetMap, Plan, UniformGrid

__all__ = ["HIP", "MGA", "Constant", "Grid", "OffsetMap", "Plan", "UniformGrid"]
