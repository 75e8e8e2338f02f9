"""Predefined simulators: the EPG trains and the bSSFP closed form."""

from .bssfp import bSSFPSimulator
from .fse import FSESimulator
from .mrf import MRFSimulator
from .truefisp import TrueFISPSimulator

__all__ = ["FSESimulator", "bSSFPSimulator", "MRFSimulator", "TrueFISPSimulator"]
