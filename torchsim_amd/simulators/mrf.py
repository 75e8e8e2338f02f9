"""Unbalanced SSFP MR fingerprinting (reference simulators/mrf.py:17-84)."""

from __future__ import annotations

import threading
from typing import Any

import numpy as np
import torch

from ..model._simulator import Simulator, SpinPhysics, _protocol_key
from ..sequence._engine import carries_derivative
from ..sequence._events import INVERSION, UNBALANCED, Inversion, PackedEvents, host_values

# one two-pulse template per protocol (everything but the flips and TRs)
_TEMPLATES: dict = {}
_TEMPLATE_LOCK = threading.Lock()


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
    """Inversion-prepared MR fingerprinting over a train of N excitations.

    Pulse k plays flip_k at phase_k; its echo is sampled, one spoiler gradient
    advances every configuration by one order, and the spins relax for the
    rest of TR_k. The returned signal is the recorded sample times i.
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
        """The packed stream, built in a handful of vectorised ops instead of
        one Python object per event (the generic packing costs ~25 ms for a
        thousand pulses; this ~1 ms), bit for bit the generic buffers.

        The structure (kinds, actions, phases, output slots) is that of the
        generic packing of the same protocol with two placeholder pulses,
        cached and tiled to N pulses; the flips are scattered into it; the
        intervals are the float32 clock differences the generic packing
        computes, evaluated with the same float32 operations in the same
        order. When the flips and/or TRs carry a graph (sequence design) the
        intervals also carry a term that is zero in value and has
        d(interval)/d(TR) = 1e-3 s/ms. Per-pulse float32 phases (a
        phase-cycled protocol, with or without a graph) are added onto the
        template's phases at the events the template shows them reach, with
        the sign it shows. Anything else (repetitions, a TI that is a tensor,
        phases of another dtype, non-float32 flips) takes the generic packing.
        """
        flip, tr = sequence.get("flip"), sequence.get("TR")
        other = [k for k, v in sequence.items() if k not in ("flip", "TR", "phases") and carries_derivative(v)]
        ti = sequence.get("TI", 0.0)
        phases = sequence.get("phases", 0.0)
        cycled = isinstance(phases, torch.Tensor)
        if (other or flip is None or tr is None or not isinstance(ti, (int, float))
                or not (isinstance(phases, (int, float)) or (cycled and phases.dtype == torch.float32
                                                              and phases.ndim <= 1))
                or int(sequence.get("repetitions", self.repetitions) or 1) > 1):
            return super()._packed(**sequence)
        angles = torch.deg2rad(torch.atleast_1d(torch.as_tensor(flip)))
        if angles.dtype != torch.float32 or angles.ndim != 1 or angles.numel() == 0:
            return super()._packed(**sequence)
        n = angles.numel()
        spacing_s = _matched(tr, angles) * 1e-3
        block = self._pulse_block(sequence)
        if block is None:
            return super()._packed(**sequence)
        head, body, turn_at, record_at, phase_sign = block

        def tiled(name: str) -> torch.Tensor:
            return torch.cat([getattr(head, name), getattr(body, name).repeat(n)])

        pulse = torch.arange(n) * 4 + 1
        output_index = tiled("output_index")
        output_index[pulse + record_at] = torch.arange(n, dtype=output_index.dtype)
        flips = tiled("flip").index_put((pulse + turn_at,), angles.to(body.flip.dtype))
        # the clock as compose() accumulates it: a Python float until the first
        # TR is added, float32 after (add.accumulate is sequential float32)
        steps = spacing_s.detach().to(torch.float32).numpy()
        cursor0 = 0.0 + float(ti) * 1e-3
        cursors = np.add.accumulate(np.concatenate([np.array([cursor0], dtype=np.float32), steps[:-1]]),
                                    dtype=np.float32)
        stamps = np.float32(1e6) * cursors
        stamps[0] = np.float32(1e6 * cursor0)
        clock = np.concatenate([np.zeros(1, dtype=np.float32), np.repeat(stamps, 4)])
        intervals = np.diff(clock, prepend=np.float32(0.0)) * np.float32(1e-6)
        duration = torch.from_numpy(intervals.astype(np.float32))
        # the zero intervals a graph reaches: only the TRs carry one, onto the
        # excitations after the first
        reached = np.zeros(1 + 4 * n, dtype=bool)
        if carries_derivative(tr):
            carried = pulse[1:] + turn_at
            live = torch.zeros(1 + 4 * n, dtype=torch.float32).index_put(
                (carried,), (spacing_s[:-1] - spacing_s[:-1].detach()).to(torch.float32))
            duration = duration + live
            reached[carried.numpy()] = intervals[carried.numpy()] == 0
        phase = tiled("phase")
        if cycled:
            # the turns as the layout computes them, onto the events they reach
            turns = torch.deg2rad(_matched(phases, angles))
            phase = phase + torch.cat([torch.zeros(1, dtype=phase.dtype),
                                       phase_sign.repeat(n) * turns.repeat_interleave(4).to(phase.dtype)])
        packed = PackedEvents(duration=duration, kind=tiled("kind"), flip=flips, phase=phase,
                              action=tiled("action"), output_index=output_index, shim_index=tiled("shim_index"),
                              saturation=tiled("saturation"), rf_frequency_hz=tiled("rf_frequency_hz"),
                              output_count=n)
        packed.__dict__["_zero_intervals"] = reached
        return packed

    def _pulse_block(self, sequence: dict) -> tuple | None:
        """(head, one pulse's four events, the excitation's and the sample's
        place in them) from the generic packing of the protocol with two
        placeholder pulses, or None where that packing is not one inversion
        then identical four-event pulses."""
        cycled = isinstance(sequence.get("phases"), torch.Tensor)
        shared = {k: v for k, v in sequence.items() if k not in ("flip", "TR") and not (cycled and k == "phases")}
        key = _protocol_key(type(self), {**shared, "per_pulse_phases": cycled})
        if key is None:
            return None
        with _TEMPLATE_LOCK:
            if key in _TEMPLATES:
                return _TEMPLATES[key]
        placeholder = {**shared, "flip": torch.ones(2, dtype=torch.float32),
                       "TR": torch.ones(2, dtype=torch.float32)}
        if cycled:
            placeholder["phases"] = torch.zeros(2, dtype=torch.float32)
        t = Simulator._packed(self, **placeholder)
        # where a pulse's phase lands in its four events, and with which sign:
        # the difference a quarter turn makes (none without per-pulse phases)
        phase_sign = torch.zeros(4, dtype=t.phase.dtype)
        if cycled:
            turned = Simulator._packed(self, **{**placeholder, "phases": torch.full((2,), 90.0)})
            quarter = torch.deg2rad(torch.full((), 90.0))
            phase_sign = ((turned.phase[1:5] - t.phase[1:5]) / quarter).detach()
            if not bool(torch.isin(phase_sign, torch.tensor([-1.0, 0.0, 1.0])).all()) or not torch.equal(
                    turned.phase[5:9] - t.phase[5:9], turned.phase[1:5] - t.phase[1:5]):
                with _TEMPLATE_LOCK:
                    _TEMPLATES[key] = None
                return None
        block = None
        if t.kind.numel() == 9 and t.output_count == 2:
            fields = ("kind", "action", "shim_index", "saturation", "rf_frequency_hz", "phase", "flip")
            periodic = all(torch.equal(getattr(t, f)[1:5], getattr(t, f)[5:9]) for f in fields)
            turning = ((t.kind[1:5] == 1) & ((t.action[1:5] & INVERSION) == 0)).nonzero().reshape(-1)
            recorded = (t.output_index[1:5] >= 0).nonzero().reshape(-1)
            if (periodic and turning.numel() == 1 and recorded.numel() == 1
                    and int(t.output_index[1 + int(recorded)]) == 0 and int(t.output_index[5 + int(recorded)]) == 1
                    and int(t.kind[0]) == 1 and bool((t.action[0] & INVERSION) != 0)):
                head = PackedEvents(*(b[:1].detach() for b in t.buffers()), output_count=0)
                body = PackedEvents(*(b[1:5].detach() for b in t.buffers()), output_count=0)
                block = (head, body, int(turning), int(recorded), phase_sign)
        with _TEMPLATE_LOCK:
            _TEMPLATES[key] = block
        return block
