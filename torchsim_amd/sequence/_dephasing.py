"""Off-resonance and a static field spread applied to the recorded samples.

Reference: sequence/_accelerators.py:434-497 (``_refuse_a_spread``,
``_analytic_dephasing``), :820-857 (``_windings_are_uniform``), :950-1165
(the unrefocused time the packing records per sample), :647-652 and :736-737
(B0 held out of the states, the factor applied to the signal),
sequence/_simulation.py:740-757 (R2' = 1000 / T2').

A gradient winds the transverse states through orders; a bulk off-resonance
turns them through a phase that grows with time. Where every stretch between
two pulses whose transverse states survive winds its orders at one steady
rate, an order stands for a fixed dephasing time, and the coherence a sample
reads has gone unrefocused for the time since the last pulse less the time
the shifts since that pulse stand for -- whatever pathway it took. The
off-resonance is then exactly a turn of the sample, ``exp(-2 pi i B0 tau)``,
and a Lorentzian spread of the field across the voxel a damping
``exp(-R2' |tau|)``: both reach the signal and every Jacobian column as one
complex factor, and B0 is held out of the states, so a train whose pulses
share an axis keeps the real-subspace kernel. Where the winding is not
uniform (a balanced train, a varying TR) B0 stays in the states and a spread
is refused: the states carry no time, so nothing could apply it.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from ._events import INVERSION, POST_SHIFT, PRE_SHIFT, RECORD, SHIFT_AFTER, SPOIL_AFTER, EventType, PackedEvents

#: Stretches are compared to this fraction of the train's span: float32
#: timestamps hold a late stretch to a few parts in ten million of where the
#: train has got to (reference sequence/_accelerators.py:825).
TIMESTAMP_RESOLUTION = 16.0 * float(np.finfo(np.float32).eps)

#: Relaxation times below this are clamped before inversion (reference
#: sequence/_simulation.py:59, the kernels' own floor).
MINIMUM_RELAXATION_TIME_MS = 1e-6


@dataclass(frozen=True)
class Winding:
    """What the stream says about dephasing: whether the analytic form holds,
    and per recorded sample (in output order) the event of the last pulse
    that turned the transverse states before it (-1: none), the sample's own
    event, and the orders wound since that pulse; ``rate`` is the pair of
    pulses and the orders of the first surviving stretch that winds, which
    fixes the time one order stands for (None if none winds)."""

    analytic: bool
    pulse: np.ndarray
    sample: np.ndarray
    shifts: np.ndarray
    rate: tuple[int, int, int] | None


def _uniform(stretches: list[tuple[int, float]], span_us: float) -> bool:
    """Every surviving stretch winds at the first one's rate (a stretch that
    winds nothing, or a train with no surviving stretch that winds, fails;
    no surviving stretch at all passes)."""
    if not stretches:
        return True
    orders, length = stretches[0]
    if orders < 1 or not length > 0.0:
        return False
    tolerance = TIMESTAMP_RESOLUTION * max(span_us, length)
    per_order = length / orders
    return all(count >= 1 and abs(span - count * per_order) <= tolerance for count, span in stretches)


def winding(packed: PackedEvents) -> Winding:
    """The stream's dephasing structure, worked out once per packed stream
    from its kinds, actions, output slots and intervals (host values)."""
    known = packed.__dict__.get("_winding")
    if known is not None:
        return known
    kind = packed.kind.detach().cpu().numpy()
    action = packed.action.detach().cpu().numpy().astype(np.int64)
    output = packed.output_index.detach().cpu().numpy()
    clock_us = np.cumsum(packed.duration.detach().cpu().numpy().astype(np.float64)) * 1e6
    n = int(packed.output_count)
    pulse = np.full(n, -1, dtype=np.int64)
    sample = np.full(n, -1, dtype=np.int64)
    shifts = np.zeros(n, dtype=np.int64)
    stretches: list[tuple[int, float]] = []
    rate = None
    orders, spoiled, turned, last = 0, False, False, -1
    for e in range(kind.size):
        a = int(action[e])
        if a & PRE_SHIFT:
            orders += 1
        if kind[e] == EventType.RF and not a & INVERSION:
            # a pulse that turns transverse magnetization closes a stretch
            span = clock_us[e] - (clock_us[last] if last >= 0 else 0.0)
            if turned and not spoiled and span > 0.0:
                stretches.append((orders, span))
                if rate is None and orders > 0:
                    rate = (last, e, orders)
            turned, orders, spoiled, last = True, 0, False, e
        elif kind[e] == EventType.ADC and a & RECORD and 0 <= output[e] < n:
            pulse[output[e]] = last
            sample[output[e]] = e
            shifts[output[e]] = orders
        if a & POST_SHIFT:
            orders += 1
        if a & SPOIL_AFTER:
            spoiled = True
        elif a & SHIFT_AFTER:
            orders += 1
    span_us = float(clock_us[-1]) if clock_us.size else 0.0
    found = Winding(analytic=_uniform(stretches, span_us), pulse=pulse, sample=sample, shifts=shifts, rate=rate)
    packed.__dict__["_winding"] = found
    return found


def unrefocused_s(packed: PackedEvents, device: torch.device | None = None) -> torch.Tensor:
    """Signed time each recorded sample has gone unrefocused, in seconds
    (float32, output order): the time since the last pulse less the shifts
    since it times the time one order stands for. Formed from the interval
    buffer with torch operations at double width, so a duration that carries
    a derivative passes it on."""
    found = winding(packed)
    duration = packed.duration
    if device is not None:
        duration = duration.to(device)
    clock = torch.cumsum(duration.to(torch.float64), 0)
    at = lambda index: torch.where(  # noqa: E731 - the clock at an event, 0 before the first
        torch.as_tensor(index >= 0, device=clock.device),
        clock[torch.as_tensor(np.maximum(index, 0), device=clock.device)],
        torch.zeros((), dtype=clock.dtype, device=clock.device))
    if found.sample.size == 0:
        return torch.zeros(0, dtype=torch.float32, device=clock.device)
    since = at(found.sample) - at(found.pulse)
    if found.rate is not None:
        first, second, orders = found.rate
        per_order = (clock[second] - clock[first]) / orders
        since = since - torch.as_tensor(found.shifts, dtype=torch.float64, device=clock.device) * per_order
    return since.to(torch.float32)


def applies(value: object) -> bool:
    """A property that carries a derivative, or differs from zero anywhere."""
    if isinstance(value, torch.Tensor):
        return bool(value.requires_grad) or bool(value.detach().ne(0).any())
    return float(value) != 0.0


def refuse_a_spread() -> None:
    """A spread on a stream whose winding is not uniform has no time to act by."""
    raise ValueError(
        "T2prime damps each sample by the time it has gone unrefocused, and this sequence's winding does not "
        "say it: its dephasing orders stand for gradient area, not elapsed time, and only a train that winds "
        "every surviving stretch at one steady rate ties the two (a balanced train keeps its coherences unwound; "
        "a varying TR winds alike in unlike times)"
    )


def r2_prime(t2_prime_ms: torch.Tensor) -> torch.Tensor:
    """The spread's rate, 1000 / T2' (T2' in ms, clamped to the kernels' floor)."""
    return 1000.0 / t2_prime_ms.clamp_min(MINIMUM_RELAXATION_TIME_MS)


@dataclass
class Dephasing:
    """The sample factor of one run: B0 (Hz) and T2' (ms) per atom (either
    may be None) and the unrefocused time per sample, all on the device."""

    b0_hz: torch.Tensor | None
    t2_prime_ms: torch.Tensor | None
    tau_s: torch.Tensor

    def factor(self) -> torch.Tensor:
        """(atoms, samples) complex64: exp(-R2' |tau|) exp(-2 pi i B0 tau)."""
        tau = self.tau_s[None, :]
        damping = None if self.t2_prime_ms is None else torch.exp(-r2_prime(self.t2_prime_ms)[:, None] * tau.abs())
        if self.b0_hz is None:
            return torch.complex(damping, torch.zeros_like(damping))
        angle = (-2.0 * torch.pi) * self.b0_hz[:, None] * tau
        return torch.polar(torch.ones_like(angle) if damping is None else damping, angle)

    def column(self, signal: torch.Tensor, name: str) -> torch.Tensor:
        """The Jacobian column along B0 or T2' of ``signal`` (which carries
        the factor already), (atoms, samples): -2 pi i tau s, and
        (1000 / T2'^2) |tau| s above the clamp (zero below it)."""
        tau = self.tau_s[None, :]
        if name == "b0_hz":
            return signal * torch.complex(torch.zeros_like(tau), (-2.0 * torch.pi) * tau)
        t2p = self.t2_prime_ms
        slope = torch.where(t2p > MINIMUM_RELAXATION_TIME_MS, 1000.0 / (t2p * t2p), torch.zeros_like(t2p))
        return signal * (slope[:, None] * tau.abs())
