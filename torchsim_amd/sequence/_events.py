"""Event vocabulary, operators, and the packed stream the kernels walk.

A sequence is a list of operators laid end to end (``compose``); each emits
events at absolute times. ``pack`` turns the events into the reference's nine
per-event buffers: the interval before each event, its kind, flip, phase,
action word and output slot (reference sequence/_accelerators.py:913-1170).

The vocabulary has to match the reference's because the kernels' action word
is shared: ``EventAction`` values are the bits the packed word carries
(reference sequence/_description.py:38-60), and the operators emit the same
events in the same order (reference sequence/_operators.py:145-534) so that a
train packs to the same buffers.
"""

from __future__ import annotations

import math

import threading
from collections.abc import Callable, Sequence
from dataclasses import dataclass
from enum import IntEnum, IntFlag
from typing import Any

import numpy as np
import torch
from torch.autograd import forward_ad

# Bits of the packed action word beyond EventAction's own (reference
# src/torchsim/_epg_cpu.cpp:36-41; sequence/_accelerators.py EXCITATION=64,
# REFOCUSING=128).
PRE_SHIFT = 1
POST_SHIFT = 2
INVERSION = 4
SPOIL_AFTER = 8
SHIFT_AFTER = 16
RECORD = 32
EXCITATION = 64
REFOCUSING = 128

_TO_US = 1e6

#: The power an ideal hard pulse deposits in a semisolid pool, per flip
#: squared, in 1/s: -pi int |w|^2 dt / |int w dt|^2 for the reference's ideal
#: pulse, a unit envelope over 1/(2 pi) s (sequence/_description.py:501-540,
#: 953-981), which is -2 pi^2. Packed for every pulse.
IDEAL_SATURATION = -2.0 * math.pi * math.pi


class EventType(IntEnum):
    """What an event is: time passing, a pulse, or a sample."""

    WAIT = 0
    RF = 1
    ADC = 2


class EventAction(IntFlag):
    """What the sequence does around an event: crushers, winding, spoiling."""

    NONE = 0
    CRUSH_BEFORE = 1
    CRUSH_AFTER = 2
    SPOIL_AFTER = 8
    SHIFT_AFTER = 16


class RfUse(IntEnum):
    """The role a pulse is tagged with."""

    UNKNOWN = 0
    EXCITATION = 1
    REFOCUSING = 2
    INVERSION = 3
    SATURATION = 4
    PREPARATION = 5
    OTHER = 6


@dataclass(frozen=True)
class SequenceEvent:
    """One event at an absolute time (microseconds)."""

    type: EventType
    timestamp_us: Any
    flip_rad: Any = 0.0
    phase_rad: Any = 0.0
    action: int = 0
    use: RfUse = RfUse.UNKNOWN
    recorded: bool = True


@dataclass(frozen=True)
class Operator:
    """One module of a sequence: the events it plays and the time it holds."""

    emit: Callable[[Any], tuple[SequenceEvent, ...]]
    duration_s: Any = 0.0


def compose(*parts: Operator | tuple[Any, Operator], start_s: Any = 0.0) -> tuple[tuple[SequenceEvent, ...], Any]:
    """Lay operators end to end; ``(offset_s, op)`` places one at ``start_s + offset``."""
    events: list[SequenceEvent] = []
    cursor = start_s
    for part in parts:
        offset, played = part if isinstance(part, tuple) else (None, part)
        when = cursor if offset is None else start_s + offset
        events.extend(played.emit(when))
        cursor = when + played.duration_s
    return tuple(events), cursor


def module(*parts: Operator | tuple[Any, Operator], duration_s: Any) -> Operator:
    """Several operators packaged as one placeable operator."""

    def emit(start_s: Any) -> tuple[SequenceEvent, ...]:
        events, _ = compose(*parts, start_s=start_s)
        return events

    return Operator(emit, duration_s)


def _rf(flip_rad: Any, phase_rad: Any, use: RfUse, action: int, duration_s: Any) -> Operator:
    def emit(start_s: Any) -> tuple[SequenceEvent, ...]:
        return (SequenceEvent(EventType.RF, _TO_US * start_s, flip_rad, phase_rad, int(action), use),)

    return Operator(emit, duration_s)


def Excitation(flip_rad: Any, phase_rad: Any = 0.0, *, duration_s: Any = 0.0, action: int = 0) -> Operator:
    """A pulse tipping magnetization into the transverse plane."""
    return _rf(flip_rad, phase_rad, RfUse.EXCITATION, action, duration_s)


def Refocusing(flip_rad: Any, phase_rad: Any = 0.0, *, crushed: bool = True, duration_s: Any = 0.0) -> Operator:
    """A refocusing pulse, between crushers unless ``crushed`` is false."""
    action = (EventAction.CRUSH_BEFORE | EventAction.CRUSH_AFTER) if crushed else EventAction.NONE
    return _rf(flip_rad, phase_rad, RfUse.REFOCUSING, int(action), duration_s)


def Inversion(*, duration_s: Any = 0.0, flip_rad: Any = torch.pi, phase_rad: Any = 0.0) -> Operator:
    """An ideal inversion: ``Mz -> -efficiency * Mz`` on every order."""
    return _rf(flip_rad, phase_rad, RfUse.INVERSION, 0, duration_s)


def Readout(phase_rad: Any = 0.0, *, duration_s: Any = 0.0, action: int = 0, recorded: bool = True) -> Operator:
    """One sample of order-zero F+, demodulated by ``phase_rad``."""

    def emit(start_s: Any) -> tuple[SequenceEvent, ...]:
        return (SequenceEvent(EventType.ADC, _TO_US * start_s, 0.0, phase_rad, int(action), recorded=recorded),)

    return Operator(emit, duration_s)


def Delay(duration_s: Any, *, action: int = 0) -> Operator:
    """Time passing, with whatever the sequence plays across it."""

    def emit(start_s: Any) -> tuple[SequenceEvent, ...]:
        return (SequenceEvent(EventType.WAIT, _TO_US * start_s, action=int(action)),)

    return Operator(emit, duration_s)


def Dephase() -> Operator:
    """An unbalanced gradient taking no time: every order winds on by one."""
    return Operator(lambda start_s: (SequenceEvent(EventType.WAIT, _TO_US * start_s, action=int(EventAction.CRUSH_AFTER)),))


def Spoil() -> Operator:
    """Ideal transverse spoiling taking no time."""
    return Operator(lambda start_s: (SequenceEvent(EventType.WAIT, _TO_US * start_s, action=int(EventAction.SPOIL_AFTER)),))


def SSFPFidReadout(phase_rad: Any = 0.0, *, duration_s: Any = 0.0) -> Operator:
    """A sample followed by one unbalanced gradient (reference _operators.py:471-489)."""
    return module(Readout(phase_rad), Dephase(), Delay(duration_s), duration_s=duration_s)


def bSSFPReadout(phase_rad: Any = 0.0, *, duration_s: Any = 0.0) -> Operator:
    """A sample the repetition rewinds after (reference _operators.py:449-468)."""
    return module(Readout(phase_rad), Delay(duration_s), duration_s=duration_s)


def SPGRReadout(phase_rad: Any = 0.0, *, duration_s: Any = 0.0) -> Operator:
    """A sample followed by ideal spoiling (reference _operators.py:521-537)."""
    return module(Readout(phase_rad), Spoil(), Delay(duration_s), duration_s=duration_s)


def FSEReadout(phase_rad: Any = 0.0, *, duration_s: Any = 0.0) -> Operator:
    """The sample at a spin echo (reference _operators.py:540-556)."""
    return module(Readout(phase_rad), Delay(duration_s), duration_s=duration_s)


@dataclass(frozen=True)
class EventOperators:
    """Which operator plays each role of a sequence."""

    excitation: Callable[..., Operator] = Excitation
    refocusing: Callable[..., Operator] = Refocusing
    inversion: Callable[..., Operator] = Inversion
    readout: Callable[..., Operator] = Readout
    delay: Callable[..., Operator] = Delay


def _uncrushed(*args: Any, **kwargs: Any) -> Operator:
    return Refocusing(*args, crushed=False, **kwargs)


#: Readout rewound after, refocusing uncrushed (reference model/_state_machine.py:275).
BALANCED = EventOperators(readout=bSSFPReadout, refocusing=_uncrushed)
#: Readout followed by one unbalanced gradient.
UNBALANCED = EventOperators(readout=SSFPFidReadout)
#: Readout followed by ideal spoiling.
SPOILED = EventOperators(readout=SPGRReadout)
#: Refocusing between crushers, sample at the echo centre.
REFOCUSED = EventOperators(readout=FSEReadout)


@dataclass
class PackedEvents:
    """The nine per-event buffers, plus what the host needs to read them."""

    duration: torch.Tensor  # f32 (E,) seconds
    kind: torch.Tensor  # i32 (E,)
    flip: torch.Tensor  # f32 (E,) rad
    phase: torch.Tensor  # f32 (E,) rad
    action: torch.Tensor  # u8 (E,)
    output_index: torch.Tensor  # i32 (E,)
    shim_index: torch.Tensor  # i32 (E,)
    saturation: torch.Tensor  # f32 (E,)
    rf_frequency_hz: torch.Tensor  # f32 (E,)
    output_count: int

    @property
    def event_count(self) -> int:
        return int(self.kind.numel())

    def buffers(self) -> list[torch.Tensor]:
        return [
            self.duration,
            self.kind,
            self.flip,
            self.phase,
            self.action,
            self.output_index,
            self.shim_index,
            self.saturation,
            self.rf_frequency_hz,
        ]

    def to(self, device: torch.device | str) -> PackedEvents:
        moved = [value.to(device) for value in self.buffers()]
        return PackedEvents(*moved, output_count=self.output_count)

    def on(self, device: torch.device) -> list[torch.Tensor]:
        """The nine buffers on ``device``, contiguous, copied there once.

        The copy is asynchronous (one pinned staging buffer, one transfer on
        the current stream); every call makes the current stream wait for it,
        so callers use the buffers on their own stream without a host wait.
        """
        if any(value.requires_grad or forward_ad.unpack_dual(value).tangent is not None for value in self.buffers()):
            return [value.to(device).contiguous() for value in self.buffers()]
        return resident(self, "events", self.buffers(), device)


class _Upload:
    """Device copies of some host tensors and the event that says they landed."""

    def __init__(self, tensors: list[torch.Tensor | None], device: torch.device) -> None:
        if torch.cuda.is_current_stream_capturing():
            # a captured copy would re-read a staging buffer that is reused,
            # and the cached device copy would live in the graph's memory
            raise RuntimeError("device buffers are uploaded outside graph capture: run the launch eagerly first")
        host = [None if t is None else t.detach().contiguous() for t in tensors]
        self.values: list[torch.Tensor | None] = []
        self.ready: torch.cuda.Event | None = None
        if any(t is not None and t.device.type != "cpu" for t in host):
            self.values = [None if t is None else t.to(device) for t in host]
            return
        offsets, total = [], 0
        for t in host:
            offsets.append(total)
            if t is not None:
                total += -(-t.numel() * t.element_size() // 256) * 256
        if total == 0:
            self.values = [None if t is None else t.to(device) for t in host]
            return
        staging = torch.empty(total, dtype=torch.uint8, pin_memory=True)
        for t, at in zip(host, offsets):
            if t is not None and t.numel():
                staging[at:at + t.numel() * t.element_size()].copy_(t.reshape(-1).view(torch.uint8))
        stream = torch.cuda.current_stream(device)
        with torch.cuda.device(device):
            landed = staging.to(device, non_blocking=True)
            self.ready = torch.cuda.Event()
            self.ready.record(stream)
        for t, at in zip(host, offsets):
            if t is None:
                self.values.append(None)
            else:
                nbytes = t.numel() * t.element_size()
                self.values.append(landed[at:at + nbytes].view(t.dtype).reshape(t.shape))

    def wait(self, device: torch.device) -> list[torch.Tensor | None]:
        ready = self.ready
        if ready is not None:
            if torch.cuda.is_current_stream_capturing():
                # a graph cannot wait on a copy made outside it
                raise RuntimeError("device buffers are still being uploaded: run the launch once, synchronize "
                                   "and run it again before capturing it")
            if ready.query():
                self.ready = None
            else:
                torch.cuda.current_stream(device).wait_event(ready)
        return self.values


_RESIDENT_LOCK = threading.Lock()


def resident(owner: Any, name: str, tensors: list[torch.Tensor | None], device: torch.device) -> list:
    """``tensors`` (host) on ``device``, uploaded once per ``owner`` and
    ``name`` (asynchronously; the current stream waits for the copy). Callers
    on several threads (a client building requests, the serving loop) share
    one upload per key."""
    cache = owner.__dict__.setdefault("_resident", {})
    key = (name, str(device))
    upload = cache.get(key)
    if upload is None:
        with _RESIDENT_LOCK:
            upload = cache.get(key)
            if upload is None:
                upload = _Upload(list(tensors), device)
                cache[key] = upload
    return upload.wait(device)


def host_values(*arrays: Any) -> tuple[Any, ...]:
    """A layout's per-pulse arrays as host NumPy arrays of the same dtype when
    no autograd graph runs through them. Indexing them then yields NumPy
    scalars, whose float32 / float64 arithmetic is the same IEEE arithmetic as
    that of 0-dim tensors of the dtype (Python numbers stay weak in both), so
    the packed buffers are bit for bit those the tensors would give, at a
    small fraction of the per-operation cost of a layout of a thousand pulses.
    """
    if any(_carries(a) for a in arrays):
        return arrays
    return tuple(a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a for a in arrays)


def _carries(value: Any) -> bool:
    """A tensor that requires grad or holds a forward-mode tangent."""
    return isinstance(value, torch.Tensor) and (
        value.requires_grad or forward_ad.unpack_dual(value).tangent is not None)


def _graphed(values: Sequence[Any]) -> bool:
    return any(_carries(value) for value in values)


def _stack(values: Sequence[Any]) -> torch.Tensor:
    """Stack per-event values at float32, keeping any autograd graph."""
    if not _graphed(values):
        return torch.from_numpy(np.array([float(value) for value in values], dtype=np.float64).astype(np.float32))
    if any(isinstance(value, torch.Tensor) for value in values):
        return torch.stack(
            [
                value.to(torch.float32).reshape(()) if isinstance(value, torch.Tensor) else torch.tensor(float(value), dtype=torch.float32)
                for value in values
            ]
        )
    return torch.tensor([float(value) for value in values], dtype=torch.float32)


def _clock(values: Sequence[Any]) -> torch.Tensor:
    """The stream's timestamps at the width the description timed them in.

    Timestamps accumulated from tensors (or the NumPy scalars of
    ``host_values``) are float32 (reference sequence/_accelerators.py:1222-1247);
    plain Python numbers stay float64.
    """
    timed = any(isinstance(value, (torch.Tensor, np.floating)) for value in values)
    if timed and not _graphed(values):
        return torch.from_numpy(np.array([float(value) for value in values], dtype=np.float64).astype(np.float32))
    if timed:
        return torch.stack(
            [
                value.to(torch.float32).reshape(()) if isinstance(value, torch.Tensor) else torch.tensor(float(value), dtype=torch.float32)
                for value in values
            ]
        )
    return torch.tensor([float(value) for value in values], dtype=torch.float64)


def pack(events: Sequence[SequenceEvent], *, repetitions: int = 1, tr_duration_us: Any = 0.0) -> PackedEvents:
    """Pack a stream played ``repetitions`` times, recording the last playing.

    Follows the reference's packing (sequence/_accelerators.py:913-1170):
    intervals are one difference over the stream's clock, converted to
    seconds at float32; inversions carry INVERSION, other pulses EXCITATION or
    REFOCUSING; recorded samples carry RECORD and a running output index.
    """
    clock: list[Any] = []
    kinds: list[int] = []
    flips: list[Any] = []
    phases: list[Any] = []
    actions: list[int] = []
    outputs: list[int] = []
    output = 0
    for repetition in range(repetitions):
        recording = repetition == repetitions - 1
        offset = tr_duration_us * repetition if repetition else 0.0
        for event in events:
            clock.append(offset + event.timestamp_us if repetition else event.timestamp_us)
            kinds.append(int(event.type))
            action = int(event.action)
            flip: Any = 0.0
            phase: Any = 0.0
            index = -1
            if event.type is EventType.RF:
                flip = event.flip_rad
                phase = event.phase_rad
                if event.use is RfUse.INVERSION:
                    action |= INVERSION
                else:
                    action |= REFOCUSING if event.use is RfUse.REFOCUSING else EXCITATION
            elif event.type is EventType.ADC:
                phase = event.phase_rad
                if recording and event.recorded:
                    action |= RECORD
                    index = output
                    output += 1
            kinds_flip = flip
            flips.append(kinds_flip)
            phases.append(phase)
            actions.append(action)
            outputs.append(index)
    count = len(kinds)
    clock_us = _clock(clock) if clock else torch.zeros(0, dtype=torch.float64)
    duration = (torch.diff(clock_us, prepend=clock_us.new_zeros(1)) * 1e-6).to(torch.float32)
    return PackedEvents(
        duration=duration.contiguous(),
        kind=torch.tensor(kinds, dtype=torch.int32),
        flip=_stack(flips) if count else torch.zeros(0),
        phase=_stack(phases) if count else torch.zeros(0),
        action=torch.tensor(actions, dtype=torch.uint8),
        output_index=torch.tensor(outputs, dtype=torch.int32),
        shim_index=torch.zeros(count, dtype=torch.int32),
        saturation=torch.tensor([IDEAL_SATURATION if k == EventType.RF else 0.0 for k in kinds],
                                dtype=torch.float32),
        rf_frequency_hz=torch.zeros(count, dtype=torch.float32),
        output_count=output,
    )
