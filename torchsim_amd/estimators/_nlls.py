"""Nonlinear least squares: a Levenberg-Marquardt fit of every voxel at once
(reference estimators/_nlls.py:18-327), on the fused normal-equation kernel."""

from __future__ import annotations

import dataclasses
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
                 initial: Mapping[str, float] | None = None, loop: GaussNewton | None = None,
                 subspace: Any = None) -> None:
        self.acquisition = acquisition
        self._subspace = subspace
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

    @property
    def subspace(self) -> Any:
        """The temporal basis fits run in, if any (reference
        estimators/_mapping.py:101-104)."""
        return self._subspace

    def fit(self, *, seed: int | None = None, samples: int = 10_000, rank: int | None = None,
            subspace: Any = None, **ranges: Any) -> NonlinearLeastSquares:
        """Name the unknowns with their ranges ``name=(low, high)`` and choose
        the start. ``rank`` fits a temporal basis of that rank to ``samples``
        signals simulated uniformly over the ranges; ``subspace`` hands one in
        (reference estimators/_mapping.py:118-255). With either, every fit
        runs on subspace coefficients."""
        if not ranges:
            raise ValueError("name at least one unknown property with its range")
        if rank is not None and subspace is not None:
            raise ValueError("give a rank to fit a basis here, or a subspace to work in one, not both")
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
        if subspace is not None:
            self._subspace = subspace
        elif rank is not None:
            self._subspace = self._fit_basis(ranges, samples, rank, generator)
        return self

    def _fit_basis(self, ranges: Mapping[str, Any], samples: int, rank: int, generator: torch.Generator) -> Any:
        """A rank-``rank`` basis of the acquisition's signals over the ranges
        (uniform draws; the reference's training-set draw is its own)."""
        from .._subspace import Subspace

        if self.acquisition is None:
            raise RuntimeError("a basis is fitted to simulated signals: give this estimator its acquisition")
        values = {}
        for name, (low, high) in ranges.items():
            drawn = low + (high - low) * torch.rand(samples, generator=generator, dtype=torch.float64)
            values[name] = drawn.to(device="cuda", dtype=torch.float32)
        return Subspace.fit(torch.as_tensor(self.acquisition.simulate(**values)).cpu(), rank)

    def map(self, volume: Any, known: Mapping[str, Any] | None = None) -> dict[str, torch.Tensor]:
        """``{name: map}`` shaped like ``volume``'s leading axes."""
        if not self.fitted:
            raise RuntimeError("the estimator has no model and starting point to fit from")
        return self._solve(volume, known, project=True)

    def from_coefficients(self, coefficients: Any, known: Mapping[str, Any] | None = None) -> dict[str, torch.Tensor]:
        """Maps from coefficients already in :attr:`subspace` (what a subspace
        reconstruction returns), without projecting them again (reference
        estimators/_mapping.py:453-498)."""
        if self._subspace is None:
            raise RuntimeError("this estimator has no subspace, so there are no coefficients to read; "
                               "state a rank when it is fitted")
        values = torch.as_tensor(coefficients)
        if values.shape[-1] != self._subspace.rank:
            raise ValueError(f"the basis has rank {self._subspace.rank}, got {values.shape[-1]} coefficients")
        return self._solve(values, known, project=False)

    def _solve(self, volume: Any, known: Mapping[str, Any] | None, *, project: bool) -> dict[str, torch.Tensor]:
        if not self.fitted:
            raise RuntimeError("the estimator has no model and starting point to fit from")
        signals = torch.as_tensor(volume)
        if signals.device.type != "cuda":
            signals = signals.cuda()
        shape = signals.shape[:-1]
        measured = signals.reshape(-1, signals.shape[-1])
        subspace = self._subspace
        if subspace is not None:
            # the basis is read by every iteration's launch: hold it where they run
            subspace = dataclasses.replace(subspace, basis=subspace.basis.to(measured.device))
            if project:
                measured = subspace.project(measured)
        operator = FusedModel(self.acquisition, self.unknown, bounds=self.bounds, known=known, subspace=subspace)
        start = operator.free(self._start.to(measured.device)[None, :].expand(measured.shape[0], -1))
        found = self.loop.minimize(operator, measured, start)
        self.iterations = found.iterations
        self.unconverged = found.unconverged
        self.cost = found.cost
        values = operator.natural(found.x)
        return {name: values[:, i].reshape(shape).to(torch.float32) for i, name in enumerate(self.unknown)}
