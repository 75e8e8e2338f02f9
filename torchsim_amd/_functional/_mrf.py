"""MR fingerprinting wrapper (reference _functional/_mrf.py:12-85)."""

from __future__ import annotations

from typing import Any

import torch

from ..simulators.mrf import MRFSimulator
from ._run import evaluated


def mrf_sim(
    flip: Any,
    TR: Any,
    T1: Any,
    T2: Any,
    diff: str | tuple[str, ...] | None = None,
    B1: Any = 1.0,
    inv_efficiency: Any = 1.0,
    M0: Any = 1.0,
    TI: float = 0.0,
    states: int = 10,
    nreps: int = 1,
    device: str | torch.device | None = None,
) -> Any:
    """SSFP MR fingerprinting: signal ``(..., len(flip))``, and with ``diff``
    its Jacobian ``(..., len(diff), len(flip))`` (a single name drops that axis).

    Flip angles in degrees, TR/TI/T1/T2 in milliseconds.
    """
    return evaluated(
        MRFSimulator(),
        diff,
        device,
        T1=T1,
        T2=T2,
        M0=M0,
        B1=B1,
        inv_efficiency=inv_efficiency,
        flip=flip,
        TR=TR,
        TI=TI,
        states=states,
        repetitions=nreps,
    )
