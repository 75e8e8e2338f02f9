"""Functional wrappers over the predefined simulators (reference _functional/)."""

from ._bssfp import bssfp_sim
from ._fse import fse_sim
from ._mrf import mrf_sim

__all__ = ["bssfp_sim", "fse_sim", "mrf_sim"]
