"""Unbalanced SSFP MR fingerprinting (reference simulators/mrf.py:17-84)."""

from __future__ import annotations

from typing import Any

import torch

from ..model._simulator import Simulator, SpinPhysics
from ..sequence._events import UNBALANCED, Inversion


def _matched(value: Any, reference: torch.Tensor) -> torch.Tensor:
    """One entry per entry of ``reference``, in its dtype (reference _array.py:149-171)."""
    held = torch.as_tensor(value).to(reference.dtype)
    if held.numel() == 1:
        return held.reshape(()).expand(reference.shape)
    if held.shape != reference.shape:
        raise ValueError(
            f"expected one value per event or a single one, got {tuple(held.shape)} against {tuple(reference.shape)}"
        )
    return held


class MRFSimulator(Simulator):
    """An inversion, then a variable flip-angle train whose readouts wind on.

    Layout per repetition k: an excitation (flip_k, phase_k), a sample, one
    unbalanced gradient, and the rest of TR_k; the returned signal is i times
    the recorded sample.
    """

    model = SpinPhysics(
        properties={
            "T1": "t1_ms",
            "T2": "t2_ms",
            "M0": "m0",
            "B1": "b1",
            "inv_efficiency": "inversion_efficiency",
        },
        operators=UNBALANCED,
    )
    states = 10
    signal_factor = 1j

    def layout(self, *, flip: Any, TR: Any, TI: Any = 0.0, phases: Any = 0.0) -> list:
        angles = torch.deg2rad(torch.atleast_1d(torch.as_tensor(flip)))
        turns = torch.deg2rad(_matched(phases, angles))
        spacing_s = _matched(TR, angles) * 1e-3
        ops = self.model.operators
        parts = [Inversion(duration_s=TI * 1e-3)]
        for index in range(angles.numel()):
            parts.append(ops.excitation(angles[index], turns[index]))
            parts.append(ops.readout(turns[index], duration_s=spacing_s[index]))
        return parts
