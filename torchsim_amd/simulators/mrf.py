"""Unbalanced SSFP MR fingerprinting (reference simulators/mrf.py:17-84)."""

from __future__ import annotations

import dataclasses
from typing import Any

import numpy as np
import torch

from ..model._simulator import Simulator, SpinPhysics
from ..sequence._events import INVERSION, UNBALANCED, Inversion, PackedEvents, host_values


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
        angles, turns, spacing_s = host_values(angles, turns, spacing_s)
        ops = self.model.operators
        parts = [Inversion(duration_s=TI * 1e-3)]
        for index in range(len(angles)):
            parts.append(ops.excitation(angles[index], turns[index]))
            parts.append(ops.readout(turns[index], duration_s=spacing_s[index]))
        return parts

    def _packed(self, **sequence: Any) -> PackedEvents:
        """The packed stream; when only the flips and/or TRs carry a graph
        (sequence design), built in a handful of vectorised ops instead of one
        autograd node per event.

        The structure (kinds, actions, phases, output slots) comes from the
        generic packing of the same protocol with placeholder values, which is
        cached; the flips are scattered into it; the intervals are the float32
        clock differences the generic packing computes, evaluated with the same
        float32 operations in the same order, plus a term that is zero in value
        and carries d(interval)/d(TR) = 1e-3 s/ms into the graph.
        """
        flip, tr = sequence.get("flip"), sequence.get("TR")
        graphed = [v for v in (flip, tr) if isinstance(v, torch.Tensor) and v.requires_grad]
        other = [k for k, v in sequence.items()
                 if k not in ("flip", "TR") and isinstance(v, torch.Tensor) and v.requires_grad]
        ti = sequence.get("TI", 0.0)
        if (not graphed or other or isinstance(ti, torch.Tensor)
                or int(sequence.get("repetitions", self.repetitions) or 1) > 1):
            return super()._packed(**sequence)
        angles = torch.deg2rad(torch.atleast_1d(torch.as_tensor(flip)))
        n = angles.numel()
        spacing_s = _matched(tr, angles) * 1e-3
        placeholder = {**sequence, "flip": torch.ones(n, dtype=angles.dtype), "TR": torch.ones(n, dtype=angles.dtype)}
        template = self.packed(**placeholder)
        turning = ((template.kind == 1) & ((template.action & INVERSION) == 0)).nonzero().reshape(-1)
        if turning.numel() != n or template.kind.numel() != 1 + 4 * n:
            return super()._packed(**sequence)
        flips = template.flip.detach().clone().index_put((turning,), angles.to(template.flip.dtype))
        # the clock as compose() accumulates it: Python float until the first
        # TR is added, float32 tensors after
        steps = spacing_s.detach().to(torch.float32).numpy()
        clock = np.zeros(1 + 4 * n, dtype=np.float32)
        cursor: Any = 0.0 + float(ti) * 1e-3
        for k in range(n):
            stamp = np.float32(1e6) * cursor if isinstance(cursor, np.float32) else np.float32(1e6 * cursor)
            clock[1 + 4 * k: 5 + 4 * k] = stamp
            cursor = np.float32(np.float32(cursor) + steps[k])
        intervals = np.diff(clock, prepend=np.float32(0.0)) * np.float32(1e-6)
        live = torch.zeros(1 + 4 * n, dtype=torch.float32).index_put(
            (turning[1:],), (spacing_s[:-1] - spacing_s[:-1].detach()).to(torch.float32))
        duration = torch.from_numpy(intervals.astype(np.float32)) + live
        return dataclasses.replace(template, duration=duration, flip=flips)
