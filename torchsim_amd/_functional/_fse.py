"""Fast spin echo as a function (reference _functional/_fse.py:12-106)."""

from __future__ import annotations

from typing import Any

from ..simulators.fse import FSESimulator
from ._run import evaluated


def fse_sim(flip: Any, ESP: Any, T1: Any, T2: Any, phases: Any = 0.0, TR: Any = 1e6, exc_flip: float = 90.0,
            exc_phase: float = 90.0, diff: Any = None, B1: Any = 1.0, M0: Any = 1.0, states: int = 10,
            device: Any = None) -> Any:
    """The refocused train sampled at every echo (refocusing flips and phases
    in degrees, ESP and TR in ms); with ``diff``, its Jacobian too."""
    return evaluated(FSESimulator(), diff, device, T1=T1, T2=T2, M0=M0, B1=B1, flip=flip, ESP=ESP, phases=phases,
                     TR=TR, exc_flip=exc_flip, exc_phase=exc_phase, states=states)
