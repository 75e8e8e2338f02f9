"""Fast spin echo wrapper (reference _functional/_fse.py)."""

from __future__ import annotations

from typing import Any

from ..simulators.fse import FSESimulator
from ._run import evaluated


def fse_sim(flip: Any, ESP: Any, T1: Any, T2: Any, diff: Any = None, B1: Any = 1.0, M0: Any = 1.0,
            phases: Any = 0.0, TR: Any = 1e6, exc_flip: float = 90.0, exc_phase: float = 90.0, states: int = 10,
            device: Any = None) -> Any:
    """Refocused train sampled at every echo; with ``diff``, its Jacobian too."""
    return evaluated(FSESimulator(), diff, device, T1=T1, T2=T2, M0=M0, B1=B1, flip=flip, ESP=ESP, phases=phases,
                     TR=TR, exc_flip=exc_flip, exc_phase=exc_phase, states=states)
