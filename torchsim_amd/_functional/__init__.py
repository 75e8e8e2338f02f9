"""Functional wrappers over the predefined simulators (reference _functional/)."""

from ._mrf import mrf_sim

__all__ = ["mrf_sim"]
