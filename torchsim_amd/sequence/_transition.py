"""Shaped, slice-selective pulses: the rotation they perform, tabulated.

Reference: sequence/_transition.py:148-224 (``compose_spinor``), :527-628
(``transition_table``), :630-743 (``ExactSliceProfile``,
``exact_slice_profile``, ``across_the_slice``); sequence/_description.py:
463-568 (a definition's complex integral, deposited power and flip angle);
the kernels' use of the table, src/torchsim/_epg_cpu.cpp:2041-2049,
2107-2160, 2273-2304, 4491-4520.

A pulse with a duration, played under a slice-select gradient, turns a spin
at slice position z (in slice thicknesses) about an axis that leaves the
transverse plane: during sample s the effective field is
(theta w_s, 0, 2 pi BW z dt_s) with w the envelope normalized to unit
complex area, so no flip angle and phase can name the rotation; its
Cayley-Klein pair (a, b) can. The pair depends on the voxel only through the
effective flip theta = flip * B1 and the position, and the RF phase phi
factors out (b -> b exp(-i phi)), so one table over (position, theta) at
zero phase serves every pulse of the shape. It is composed sample by sample
in FP64 (each sample the SU(2) element of its effective field), referenced
to the middle of the pulse (the turn the gradient leaves either side belongs
to the gradient and is rewound), and stored with its slope along theta
(forward mode of the same composition) for a cubic Hermite read.

A voxel is simulated at every tabulated position and its signal is the mean
over them.
"""

from __future__ import annotations

import functools
import math
from dataclasses import dataclass, field

import numpy as np
import torch

__all__ = ["ExactSliceProfile", "RfPulse", "TransitionTable", "compose_spinor", "exact_slice_profile",
           "sinc_pulse", "transition_table"]


@dataclass(frozen=True)
class RfPulse:
    """A shaped pulse: its complex envelope sampled every ``raster_s``
    seconds (any scale: the flip the event names sets the amplitude) and the
    bandwidth the slice-select gradient maps one slice thickness onto (0 for
    a pulse played without slice selection)."""

    envelope: np.ndarray = field(compare=False)
    raster_s: float = 1e-6
    bandwidth_hz: float = 0.0

    def __post_init__(self) -> None:
        object.__setattr__(self, "envelope", np.asarray(self.envelope, dtype=np.complex128).reshape(-1))
        if self.envelope.size < 1:
            raise ValueError("a pulse needs at least one envelope sample")

    @property
    def key(self) -> tuple:
        return (self.envelope.tobytes(), float(self.raster_s), float(self.bandwidth_hz))

    @property
    def duration_s(self) -> float:
        return self.envelope.size * self.raster_s

    def area(self) -> complex:
        """The envelope's complex time integral (samples centred in their
        dwell, trapezoid)."""
        if self.envelope.size < 2:
            return complex(self.envelope[0] * self.raster_s)
        times = (np.arange(self.envelope.size) + 0.5) * self.raster_s
        return complex(np.trapezoid(self.envelope, times))

    def integral_phase(self) -> float:
        """The phase the pulse's own area carries, added to every event's."""
        return float(np.angle(self.area()))

    def saturation(self) -> float:
        """Power deposited in a semisolid pool per flip squared, 1/s:
        -pi int |w|^2 dt / |int w dt|^2."""
        if self.envelope.size < 2:
            return 0.0
        times = (np.arange(self.envelope.size) + 0.5) * self.raster_s
        power = float(np.trapezoid(np.abs(self.envelope) ** 2, times))
        area = abs(self.area())
        return 0.0 if area == 0.0 else -math.pi * power / (area * area)

    def selective(self) -> bool:
        """Whether a flip and a phase cannot name the rotation: a slice-select
        bandwidth, or an envelope whose phase moves."""
        if self.bandwidth_hz != 0.0:
            return True
        nonzero = self.envelope[np.abs(self.envelope) > 0]
        return nonzero.size > 0 and bool(np.ptp(np.unwrap(np.angle(nonzero))) > 1e-12)


def sinc_pulse(duration_s: float = 2e-3, time_bandwidth: float = 4.0, *, raster_s: float = 1e-6,
               window: str = "hann") -> RfPulse:
    """A windowed sinc of ``time_bandwidth`` zero crossings across its
    duration (bandwidth time_bandwidth / duration), slice-selective."""
    count = max(2, int(round(duration_s / raster_s)))
    t = (np.arange(count) + 0.5) / count - 0.5  # in durations, centred
    shape = np.sinc(time_bandwidth * t)
    if window == "hann":
        shape = shape * (0.5 + 0.5 * np.cos(2 * np.pi * t))
    return RfPulse(envelope=shape.astype(np.complex128), raster_s=raster_s,
                   bandwidth_hz=time_bandwidth / (count * raster_s))


def compose_spinor(drive: torch.Tensor, turn_z: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """The Cayley-Klein pair of a sequence of rotations: sample s turns
    right-handedly about (Re drive_s, Im drive_s, turn_z) by that vector's
    length (drive (samples, ...), turn_z broadcast against each sample).
    A spin from +z ends with Mxy = -2 conj(a b), Mz = |a|^2 - |b|^2; a real
    drive of area t on resonance gives (cos(t/2), -i sin(t/2))."""
    shape = torch.broadcast_shapes(drive.shape[1:], turn_z.shape)
    a = torch.ones(shape, dtype=torch.complex128, device=turn_z.device)
    b = torch.zeros_like(a)
    for s in range(drive.shape[0]):
        gx = drive[s].real.expand(shape)
        gy = drive[s].imag.expand(shape)
        gz = turn_z.expand(shape)
        q = gx * gx + gy * gy + gz * gz
        small = q <= 1e-18
        length = torch.sqrt(torch.where(small, torch.ones_like(q), q))
        # cos(length/2) and sin(length/2)/length as series in q near zero,
        # where the square root has no derivative to give
        c = torch.where(small, 1.0 - q / 8.0 + q * q / 384.0, torch.cos(0.5 * length))
        k = torch.where(small, 0.5 - q / 48.0 + q * q / 3840.0, torch.sin(0.5 * length) / length)
        sa = c - 1j * gz * k
        sb = -1j * (gx - 1j * gy) * k
        a, b = sa * a - sb * b.conj(), sb * a.conj() + sa * b
    return a, b


@dataclass(frozen=True)
class TransitionTable:
    """The pair and its slope along the effective flip, (positions, bins)
    complex64 each, knots from 0 to ``theta_max``."""

    a: torch.Tensor
    b: torch.Tensor
    slope_a: torch.Tensor
    slope_b: torch.Tensor
    theta_max: float

    @property
    def points(self) -> int:
        return int(self.a.shape[0])

    @property
    def bins(self) -> int:
        return int(self.a.shape[1])

    @property
    def step(self) -> float:
        return self.theta_max / (self.bins - 1)

    def packed(self, device: torch.device | str | None = None) -> torch.Tensor:
        """(positions, bins, 8) float32: a, b, da, db as (real, imag) pairs."""
        parts = (self.a.real, self.a.imag, self.b.real, self.b.imag, self.slope_a.real, self.slope_a.imag,
                 self.slope_b.real, self.slope_b.imag)
        return torch.stack(parts, -1).to(device=device, dtype=torch.float32).contiguous()


@functools.lru_cache(maxsize=16)
def _table(key: tuple, positions: tuple, bins: int, theta_max: float) -> TransitionTable:
    envelope = np.frombuffer(key[0], dtype=np.complex128)
    raster, bandwidth = key[1], key[2]
    drive = torch.tensor(envelope * raster, dtype=torch.complex128)
    weight = drive / drive.sum()  # unit complex area: on resonance theta about the phase-0 axis
    z = torch.tensor(positions, dtype=torch.float64)
    rate = 2.0 * math.pi * bandwidth * z  # rad/s at each position
    carried = rate * (envelope.size * raster)
    theta = torch.linspace(0.0, theta_max, bins, dtype=torch.float64)

    def integrate(flip: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        turn = (rate * raster)[:, None].expand(z.numel(), flip.numel())
        a, b = compose_spinor(weight[:, None, None] * flip[None, None, :], turn)
        return a * torch.exp(0.5j * carried)[:, None], b

    values, slopes = torch.func.jvp(integrate, (theta,), (torch.ones_like(theta),))
    return TransitionTable(a=values[0].to(torch.complex64), b=values[1].to(torch.complex64),
                           slope_a=slopes[0].to(torch.complex64), slope_b=slopes[1].to(torch.complex64),
                           theta_max=float(theta_max))


def transition_table(pulse: RfPulse, positions: torch.Tensor | np.ndarray, *, bins: int = 64,
                     theta_max: float = 2.0 * math.pi) -> TransitionTable:
    """The table of ``pulse`` over slice ``positions`` (slice thicknesses)
    and ``bins`` effective flips in [0, theta_max] (memoized).

    Raises
    ------
    ValueError
        If fewer than two flip knots, or a pulse that integrates to nothing.
    """
    if bins < 2:
        raise ValueError(f"a Hermite table needs at least two knots, got {bins}")
    if abs(pulse.envelope.sum()) <= 1e-6 * np.abs(pulse.envelope).sum():
        raise ValueError("the pulse integrates to nothing, so it has no flip angle to tabulate against")
    z = tuple(float(v) for v in np.asarray(positions, dtype=np.float64).reshape(-1))
    return _table(pulse.key, z, int(bins), float(theta_max))


@dataclass(frozen=True)
class ExactSliceProfile:
    """Where across the slice a shaped pulse is worked out: ``points``
    positions (or the positions themselves) spanning ``extent`` slice
    thicknesses, ``bins`` knots along the flip up to ``theta_max``."""

    points: int | tuple = 21
    extent: float = 2.0
    bins: int = 64
    theta_max: float = 2.0 * math.pi

    def positions(self) -> np.ndarray:
        if isinstance(self.points, int):
            if self.points < 1:
                raise ValueError(f"a slice needs at least one position, got {self.points}")
            if self.points == 1:
                return np.zeros(1)
            return np.linspace(-0.5 * self.extent, 0.5 * self.extent, self.points)
        return np.asarray(self.points, dtype=np.float64).reshape(-1)


def exact_slice_profile(points: int | tuple = 21, *, extent: float = 2.0, bins: int = 64,
                        theta_max: float = 2.0 * math.pi) -> ExactSliceProfile:
    """Ask for a shaped pulse to be integrated at ``points`` slice positions
    (reference sequence/_transition.py:673-711)."""
    held = points if isinstance(points, int) else tuple(float(v) for v in np.asarray(points).reshape(-1))
    return ExactSliceProfile(points=held, extent=extent, bins=bins, theta_max=theta_max)


def across_the_slice(asked: object) -> ExactSliceProfile | None:
    """A position count or a profile (reference sequence/_transition.py:714-743).

    Raises
    ------
    TypeError
        For anything else: an array of flip scalings is not a slice profile.
    """
    if asked is None or isinstance(asked, ExactSliceProfile):
        return asked
    if isinstance(asked, bool) or not isinstance(asked, int):
        raise TypeError(f"across_slice is a number of slice positions or exact_slice_profile(...), got "
                        f"{type(asked).__name__}: a slice profile is the pulse's Bloch response, not a scaling of "
                        "its flip")
    return exact_slice_profile(asked)
