"""Estimators over the fused kernels."""

from ._nlls import NonlinearLeastSquares

__all__ = ["NonlinearLeastSquares"]
