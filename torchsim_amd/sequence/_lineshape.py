"""The semisolid pool's absorption lineshape, tabulated over the pulse offset.

Reference: sequence/_lineshape.py:1-300 (``lineshape_table``,
``lineshape_reaching``), sequence/_simulation.py:80-94 (the table sized to
the largest pulse offset plus the largest voxel off-resonance), and the
kernels' cubic Hermite read (src/torchsim/_epg_cpu.cpp:1888-1910).

RF saturates the bound pool at a rate proportional to the pulse's power and
to the super-Lorentzian lineshape at the offset it is played at (Morrison and
Henkelman, MRM 33, 1995):

    G(df) = int_0^1 sqrt(2/pi) T2b / |3u^2 - 1| exp(-2 (2 pi df T2b / (3u^2 - 1))^2) du.

It is even in df and depends on nothing per voxel but the offset (T2b is a
model constant, 12 us), so one table over |df| serves every atom: knots of
value and slope at a fixed spacing, read by cubic Hermite. The integral is
taken as the grid spacing times the sum over a uniform orientation grid (the pole at
3u^2 = 1 masked: its Gaussian closes faster than the pole opens), the slope
by differentiating under the same integral. Near resonance the model
diverges; below a cutoff (snapped to a knot) the knots are filled by a cubic
spline through the integral's values between the cutoff and twice it,
mirrored about resonance so the fill stays even. The cutoff knot keeps the
integral's own slope, so the Hermite read is C1 across the join.
"""

from __future__ import annotations

import functools
import math
from dataclasses import dataclass

import numpy as np
import torch

__all__ = ["BOUND_T2_S", "Lineshape", "lineshape_reaching", "lineshape_table"]

#: T2 of the semisolid compartment, seconds (white matter's conventional value).
BOUND_T2_S = 12e-6
#: Reach and knot count of the default table: 33 kHz over 128 knots.
OFFSET_MAX_HZ = 33e3
OFFSET_BINS = 128


@dataclass(frozen=True)
class Lineshape:
    """Knots of the lineshape (seconds) and its slope in the offset
    (seconds per Hz), at ``step`` Hz spacing from resonance."""

    values: np.ndarray
    slopes: np.ndarray
    step: float
    cutoff_hz: float

    @property
    def bins(self) -> int:
        return int(self.values.size)

    @property
    def offset_max_hz(self) -> float:
        return self.step * (self.bins - 1)

    def packed(self, device: torch.device | str | None = None) -> torch.Tensor:
        """(bins, 2) float32, value then slope per knot, as the kernels read it."""
        return torch.tensor(np.stack((self.values, self.slopes), -1), dtype=torch.float32, device=device)

    def at(self, offset_hz: np.ndarray) -> np.ndarray:
        """The cubic Hermite read at |offset| (clamped to the table), FP64."""
        scaled = np.clip(np.abs(np.asarray(offset_hz, dtype=np.float64)) / self.step, 0.0, self.bins - 1.0)
        lower = np.minimum(np.floor(scaled), self.bins - 2.0)
        u = scaled - lower
        i = lower.astype(np.int64)
        h00, h10 = 2 * u**3 - 3 * u**2 + 1, u**3 - 2 * u**2 + u
        h01, h11 = -2 * u**3 + 3 * u**2, u**3 - u**2
        return (h00 * self.values[i] + h10 * self.step * self.slopes[i] + h01 * self.values[i + 1]
                + h11 * self.step * self.slopes[i + 1])


def super_lorentzian(offset_hz: np.ndarray, bound_t2_s: float = BOUND_T2_S,
                     quadrature: int = 20000) -> tuple[np.ndarray, np.ndarray]:
    """G(df) and dG/ddf on ``quadrature`` uniform points of u in [0, 1],
    FP64."""
    u = np.linspace(0.0, 1.0, quadrature)
    h = 1.0 / (quadrature - 1)
    d = 3.0 * u * u - 1.0
    keep = np.abs(d) > 1e-12
    d = np.where(keep, d, 1.0)
    # every grid point weighted by the spacing, the reference's rule (kept so
    # that the tables agree to rounding)
    weight = np.full(quadrature, h)
    scale = 2.0 * np.pi * bound_t2_s / d
    x = np.asarray(offset_hz, dtype=np.float64)[:, None] * scale[None, :]
    term = np.where(keep, np.sqrt(2.0 / np.pi) * bound_t2_s / np.abs(d) * np.exp(-2.0 * x * x), 0.0)
    return term @ weight, (term * (-4.0 * x * scale)) @ weight


@functools.lru_cache(maxsize=8)
def lineshape_table(*, bound_t2_s: float = BOUND_T2_S, offset_max_hz: float = OFFSET_MAX_HZ,
                    bins: int = OFFSET_BINS, cutoff_hz: float = 1e3, quadrature: int = 20000) -> Lineshape:
    """The tabulated lineshape (see the module docstring).

    Raises
    ------
    ValueError
        If fewer than two knots, a cutoff outside the table, or fewer than
        two knots between the cutoff and twice it.
    """
    from scipy.interpolate import InterpolatedUnivariateSpline

    if bins < 2:
        raise ValueError(f"a Hermite table needs at least two knots, got {bins}")
    if not 0.0 < cutoff_hz < offset_max_hz:
        raise ValueError(f"the cutoff {cutoff_hz} Hz must lie inside the table's {offset_max_hz} Hz")
    knots = np.linspace(0.0, offset_max_hz, bins)
    values, slopes = super_lorentzian(knots, bound_t2_s, quadrature)
    step = offset_max_hz / (bins - 1)
    edge_index = max(1, min(bins - 2, int(round(cutoff_hz / step))))
    edge = float(knots[edge_index])
    band = knots[(knots >= edge) & (knots < 2.0 * edge)]
    if band.size < 2:
        raise ValueError(f"between {edge:.0f} and {2 * edge:.0f} Hz the table has {band.size} knots; the fill "
                         "needs two (more bins or a smaller cutoff)")
    sampled, _ = super_lorentzian(band, bound_t2_s, quadrature)
    mirror = InterpolatedUnivariateSpline(np.concatenate((-band[::-1], band)),
                                          np.concatenate((sampled[::-1], sampled)), k=min(3, 2 * band.size - 1))
    inside = knots < edge
    values[inside] = mirror(knots[inside])
    slopes[inside] = mirror.derivative()(knots[inside])
    return Lineshape(values=values, slopes=slopes, step=step, cutoff_hz=edge)


def lineshape_reaching(offset_hz: float) -> Lineshape:
    """A table reaching at least ``offset_hz``: whole multiples of the
    default's reach at the default's knot spacing (so similar runs share one)."""
    blocks = max(1, math.ceil(abs(float(offset_hz)) / OFFSET_MAX_HZ))
    return lineshape_table(offset_max_hz=blocks * OFFSET_MAX_HZ, bins=blocks * (OFFSET_BINS - 1) + 1)
