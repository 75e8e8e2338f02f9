"""Predefined EPG simulators."""

from .mrf import MRFSimulator

__all__ = ["MRFSimulator"]
