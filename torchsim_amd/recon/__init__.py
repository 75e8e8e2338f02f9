"""Gauss-Newton / Levenberg-Marquardt over fused per-voxel normal equations."""

from ._bounds import to_free, to_natural, widen
from ._gauss_newton import GaussNewton, Schedule, Solution, TrustRegion, solve_damped
from ._operator import FusedModel

__all__ = ["FusedModel", "GaussNewton", "Schedule", "Solution", "TrustRegion", "solve_damped", "to_free",
           "to_natural", "widen"]
