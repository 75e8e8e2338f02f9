"""Nonlinear least squares: a Levenberg-Marquardt fit of every voxel at once
(reference estimators/_nlls.py:18-327), on the fused normal-equation kernel."""

from __future__ import annotations

from collections.abc import Mapping
from typing import Any

import torch

from ..recon import FusedModel, GaussNewton, TrustRegion


class NonlinearLeastSquares:
    """Fit the properties named at :meth:`fit` to every voxel of a volume.

    ``acquisition`` is the simulator being inverted (everything that is
    neither unknown nor measured is fixed on it); ``bounds`` maps names to
    ``(low, high)``; ``initial`` the start (else the median of a uniform
    draw over the ranges given to :meth:`fit`); ``loop`` the solve (default
    Levenberg-Marquardt, twenty steps).
    """

    def __init__(self, acquisition: Any = None, *, bounds: Mapping[str, Any] | None = None,
                 initial: Mapping[str, float] | None = None, loop: GaussNewton | None = None) -> None:
        self.acquisition = acquisition
        self.bounds = dict(bounds or {})
        self.initial = dict(initial or {})
        self.loop = loop if loop is not None else GaussNewton(TrustRegion(), max_iterations=20)
        self.unknown: tuple[str, ...] = ()
        self._start: torch.Tensor | None = None
        self.iterations = 0
        self.unconverged = 0
        self.cost: torch.Tensor | None = None

    @property
    def fitted(self) -> bool:
        return self.acquisition is not None and self._start is not None

    def fit(self, *, seed: int | None = None, samples: int = 10_000, **ranges: Any) -> NonlinearLeastSquares:
        """Name the unknowns with their ranges ``name=(low, high)`` and choose the start."""
        if not ranges:
            raise ValueError("name at least one unknown property with its range")
        generator = torch.Generator().manual_seed(0 if seed is None else seed)
        start = []
        for name, value in ranges.items():
            if name in self.initial:
                start.append(float(self.initial[name]))
                continue
            low, high = value
            drawn = low + (high - low) * torch.rand(samples, generator=generator, dtype=torch.float64)
            start.append(float(drawn.median()))
        self.unknown = tuple(ranges)
        self._start = torch.tensor(start, dtype=torch.float64)
        return self

    def map(self, volume: Any, known: Mapping[str, Any] | None = None) -> dict[str, torch.Tensor]:
        """``{name: map}`` shaped like ``volume``'s leading axes."""
        if not self.fitted:
            raise RuntimeError("the estimator has no model and starting point to fit from")
        signals = torch.as_tensor(volume)
        if signals.device.type != "cuda":
            signals = signals.cuda()
        shape = signals.shape[:-1]
        measured = signals.reshape(-1, signals.shape[-1])
        operator = FusedModel(self.acquisition, self.unknown, bounds=self.bounds, known=known)
        start = operator.free(self._start.to(measured.device)[None, :].expand(measured.shape[0], -1))
        found = self.loop.minimize(operator, measured, start)
        self.iterations = found.iterations
        self.unconverged = found.unconverged
        self.cost = found.cost
        values = operator.natural(found.x)
        return {name: values[:, i].reshape(shape).to(torch.float32) for i, name in enumerate(self.unknown)}
