"""Balanced SSFP steady state as a function (reference _functional/_bssfp.py:12-76)."""

from __future__ import annotations

from typing import Any

from ..simulators.bssfp import bSSFPSimulator
from ._run import evaluated


def bssfp_sim(flip: Any, TE: Any, TR: float, T1: Any, T2: Any, diff: Any = None, phase_inc: float = 180.0,
              B0: Any = 0.0, chemshift: Any = 0.0, M0: Any = 1.0, device: Any = None) -> Any:
    """The bSSFP steady state at TE for every flip (degrees; TE, TR, T1, T2
    in ms; B0, chemshift in Hz; phase_inc in degrees): shape
    ``(..., len(flip))``, and with ``diff`` its derivative
    ``(..., len(diff), len(flip))`` (a single name drops that axis)."""
    return evaluated(bSSFPSimulator(), diff, device, T1=T1, T2=T2, M0=M0, B0=B0, chemshift=chemshift, flip=flip,
                     TR=TR, TE=TE, phase_inc=phase_inc)
