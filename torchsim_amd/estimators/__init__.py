"""Estimators over the fused kernels."""

from ._dictionary import DictionaryMatcher, MatchResult
from ._nlls import NonlinearLeastSquares

__all__ = ["DictionaryMatcher", "MatchResult", "NonlinearLeastSquares"]
