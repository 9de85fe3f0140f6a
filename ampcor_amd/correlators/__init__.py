"""Correlation strategies and the plan that feeds them."""

from .hip import HIP
from .mga import MGA
from .plan import Constant, Field, Grid, OffsetMap, Plan, UniformGrid

__all__ = ["HIP", "MGA", "Constant", "Field", "Grid", "OffsetMap", "Plan", "UniformGrid"]
