"""Balanced SSFP (reference simulators/truefisp.py:16-169), instantaneous pulses."""

from __future__ import annotations

from typing import Any

import torch

from ..model._simulator import Simulator, SpinPhysics
from ..sequence._events import Delay, Excitation, Inversion, Readout, Spoil, host_values, module


class TrueFISPSimulator(Simulator):
    """A balanced train from equilibrium, pulses alternating in phase, each
    sample demodulated by its own pulse's phase; an alpha/2 pulse of opposite
    phase TR/2 ahead of the train; optionally an ideal inversion first.

    Pulses are instantaneous rotations (``pulse_duration=0``); a pulse played
    sample by sample is not carried.
    """

    model = SpinPhysics(properties={"T1": "t1_ms", "T2": "t2_ms", "M0": "m0", "B1": "b1", "B0": "b0_hz"})
    states = 1
    signal_factor = 1.0

    def layout(self, *, flip: Any, TR: float, nshots: int, TE: float | None = None, pulse_duration: float = 0.0,
               half_alpha: bool = True, preparation: float | None = None, inversion: str | None = None,
               spoiler: float = 0.0) -> list:
        if pulse_duration != 0:
            raise NotImplementedError("pulses are played as instantaneous rotations (pulse_duration=0)")
        repetition_s = 1e-3 * TR
        echo_s = repetition_s / 2 if TE is None else 1e-3 * TE
        if echo_s < 0 or echo_s > repetition_s:
            raise ValueError("TE is measured from the pulse and ends before the next one begins")
        angles = torch.deg2rad(torch.as_tensor(flip, dtype=torch.float64))
        if angles.numel() == 1:
            angles = angles.reshape(1).expand(nshots)
        if angles.numel() != nshots:
            raise ValueError("flip is a scalar or one value per shot")
        (angles,) = host_values(angles)
        parts: list = []
        if inversion == "ideal":
            parts.append(Inversion())
            if spoiler > 0:
                parts += [Delay(1e-3 * spoiler), Spoil()]
        elif inversion is not None:
            raise NotImplementedError("inversion is None or 'ideal'")
        if half_alpha:
            spacing_s = repetition_s / 2 if preparation is None else 1e-3 * preparation
            parts.append(Excitation(angles[0] / 2, torch.pi))
            if spacing_s > 0:
                parts.append(Delay(spacing_s))
        for shot in range(nshots):
            turn = torch.pi * (shot % 2)
            parts.append(module(Excitation(angles[shot], turn), (echo_s, Readout(turn)), duration_s=repetition_s))
        return parts
