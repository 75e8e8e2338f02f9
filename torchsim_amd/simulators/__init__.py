"""Predefined EPG simulators."""

from .fse import FSESimulator
from .mrf import MRFSimulator
from .truefisp import TrueFISPSimulator

__all__ = ["FSESimulator", "MRFSimulator", "TrueFISPSimulator"]
