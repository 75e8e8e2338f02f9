"""Balanced SSFP wrapper over the EPG TrueFISP train."""

from __future__ import annotations

from typing import Any

from ..simulators.truefisp import TrueFISPSimulator
from ._run import evaluated


def bssfp_sim(flip: Any, TR: float, nshots: int, T1: Any, T2: Any, diff: Any = None, M0: Any = 1.0, B1: Any = 1.0,
              B0: Any = 0.0, half_alpha: bool = True, inversion: str | None = None, states: int = 1,
              device: Any = None) -> Any:
    """bSSFP as a balanced EPG train (instantaneous pulses)."""
    return evaluated(TrueFISPSimulator(), diff, device, T1=T1, T2=T2, M0=M0, B1=B1, B0=B0, flip=flip, TR=TR,
                     nshots=nshots, half_alpha=half_alpha, inversion=inversion, states=states)
