"""Levenberg-Marquardt over every voxel at once, on fused normal equations.

Mirrors the reference's loop (recon/_gauss_newton.py:365-415 TrustRegion and
Schedule, :474-651 GaussNewton.minimize, :759-793 retirement): each voxel
takes its step, carries its own damping (Nielsen update), accepts or rejects
on its own gain ratio, and drops out once its gradient is flat or its step is
short, so late iterations cost only what is left.

What differs is below the loop. The reference evaluates the model and its
Jacobian (one forward and one forward-mode pass per parameter) and solves a
stacked least-squares system per voxel. Here one gfx950 launch per iteration
accumulates, per voxel and in FP64, the cost, J^T r and J^T J while it walks
the event stream -- the Jacobian never reaches memory -- and a second launch
solves the damped 2-4 parameter system in FP64, one thread per voxel.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import torch

from .. import _native

_TINY = torch.finfo(torch.float32).tiny


@dataclass
class Solution:
    """What a solve returned: the free variables, cost per step, mean damping
    per step, steps taken, voxels still moving when the loop stopped."""

    x: torch.Tensor
    cost: torch.Tensor
    damping: torch.Tensor
    iterations: int
    unconverged: int


@dataclass
class Schedule:
    """One damping for the whole volume, lowered geometrically; every step taken."""

    initial: float = 1.0
    factor: float = 1.0 / 3.0
    minimum: float = 1e-3
    anchored: bool = True
    per_voxel: bool = False

    def __post_init__(self) -> None:
        if self.initial <= 0.0:
            raise ValueError(f"initial must be positive, got {self.initial}")
        if not 0.0 < self.factor < 1.0:
            raise ValueError(f"factor must lower the damping, got {self.factor}")
        if self.minimum <= 0.0:
            raise ValueError(f"minimum must be positive, got {self.minimum}")

    def begin(self, curvature: torch.Tensor) -> torch.Tensor:
        return torch.full_like(curvature, self.initial)

    def judge(self, damping, gain, singular):
        return torch.ones_like(gain, dtype=torch.bool), (damping * self.factor).clamp_min(self.minimum)


@dataclass
class TrustRegion:
    """Levenberg-Marquardt damping per voxel with the Nielsen update."""

    tau: float = 1e-2
    anchored: bool = False
    per_voxel: bool = True
    _rising: torch.Tensor | None = field(default=None, repr=False, init=False)

    def __post_init__(self) -> None:
        if self.tau <= 0.0:
            raise ValueError(f"tau must be positive, got {self.tau}")

    def begin(self, curvature: torch.Tensor) -> torch.Tensor:
        damping = self.tau * curvature
        self._rising = torch.full_like(damping, 2.0)
        return damping

    def judge(self, damping, gain, singular):
        better = (gain > 0) & ~singular
        rising = self._rising
        lowered = damping * (1.0 - (2.0 * gain - 1.0).pow(3)).clamp_min(1.0 / 3.0)
        self._rising = torch.where(better, torch.full_like(rising, 2.0), rising * 2.0)
        return better, torch.where(better, lowered, damping * rising)

    def retire(self, kept: torch.Tensor) -> None:
        """Keep the damping growth of the voxels ``kept`` (indices) only."""
        self._rising = self._rising.index_select(0, kept)


def solve_damped(normal: torch.Tensor, scale: torch.Tensor | None, damping: torch.Tensor | None, columns: int,
                 *, step: bool = True, offset: torch.Tensor | None = None) -> tuple[torch.Tensor | None, torch.Tensor | None, torch.Tensor, torch.Tensor]:
    """(step, predicted, gradient, curvature) of every voxel, FP64 on the device."""
    voxels = normal.shape[0]
    device = normal.device
    lib = _native.load()
    out_step = torch.empty((voxels, columns), dtype=torch.float64, device=device) if step else None
    predicted = torch.empty(voxels, dtype=torch.float64, device=device) if step else None
    gradient = torch.empty((voxels, columns), dtype=torch.float64, device=device)
    curvature = torch.empty(voxels, dtype=torch.float64, device=device)
    if voxels:
        sc = None if scale is None else scale.to(torch.float64).contiguous()
        dm = None if damping is None else damping.to(torch.float64).expand(voxels).contiguous()
        of = None if offset is None else offset.to(torch.float64).contiguous()
        with torch.cuda.device(device):
            status = lib.tsim_solve_damped(
                normal.data_ptr(), None if sc is None else sc.data_ptr(), None if dm is None else dm.data_ptr(),
                None if of is None else of.data_ptr(), voxels, columns, None if out_step is None else out_step.data_ptr(),
                None if predicted is None else predicted.data_ptr(), gradient.data_ptr(), curvature.data_ptr(),
                _native.stream_of(device),
            )
        _native.check(status)
    return out_step, predicted, gradient, curvature


class GaussNewton:
    """Solve a per-voxel nonlinear least-squares problem by repeated linearization.

    ``operator`` provides ``normal(x, data, live) -> (voxels, stride) FP64`` (cost,
    J^T r, J^T J in the free variables' natural columns; ``data`` holds every
    voxel's measured row and ``live`` the voxels still iterating, so retiring
    voxels never copies the data), ``scale(x)`` (the chain-rule factor
    d natural / d free) and ``columns``.
    """

    def __init__(self, damping: Any = None, *, max_iterations: int = 8, gradient_tolerance: float = 1e-8,
                 step_tolerance: float = 1e-8) -> None:
        if max_iterations < 1:
            raise ValueError(f"max_iterations must be positive, got {max_iterations}")
        self.damping = damping if damping is not None else Schedule()
        self.max_iterations = int(max_iterations)
        self.gradient_tolerance = gradient_tolerance
        self.step_tolerance = step_tolerance

    @torch.no_grad()
    def minimize(self, operator: Any, measured: torch.Tensor, initial: torch.Tensor) -> Solution:
        shape = initial.shape[:-1]
        start = initial.reshape(-1, initial.shape[-1]).to(torch.float64)
        data = measured.reshape(start.shape[0], -1)
        columns = start.shape[-1]
        retiring = getattr(self.damping, "per_voxel", False)
        answer = start.clone()
        live = torch.arange(start.shape[0], device=start.device)
        x = start.clone()
        normal = operator.normal(x, data, live)
        cost = normal[:, 0].clone()
        _, _, gradient, curvature = solve_damped(normal, operator.scale(x), None, columns, step=False)
        damping = self.damping.begin(curvature)
        history = [cost.sum()]
        weights = []
        iterations = 0
        for _ in range(self.max_iterations):
            flat = gradient.abs().amax(-1) < self.gradient_tolerance
            if retiring:
                keep = ~flat
                x, live, data, normal, cost, damping, gradient = self._retire(keep, answer, live, x, data, normal,
                                                                              cost, damping, gradient)
                if x.shape[0] == 0:
                    break
            elif bool(flat.all()):
                break
            iterations += 1
            weights.append(damping.detach().reshape(-1).mean())
            scale = operator.scale(x)
            offset = None
            if getattr(self.damping, "anchored", False):
                # damping pulls the step toward the start: (H + a I) d = -g + a (x0 - x)
                offset = damping.reshape(-1, 1) * (start[live] - x)
            step, predicted, gradient, _ = solve_damped(normal, scale, damping, columns, offset=offset)
            walked = step.norm(dim=-1)
            short = walked < self.step_tolerance * (x.norm(dim=-1) + self.step_tolerance)
            singular = ~torch.isfinite(step).all(-1)
            candidate = torch.where(singular[:, None], x, x + step)
            trial = operator.normal(candidate, data, live)
            trial_cost = trial[:, 0]
            gain = torch.where(predicted > 0, (cost - trial_cost) / predicted.clamp_min(_TINY),
                               torch.zeros_like(predicted))
            better, damping = self.damping.judge(damping, gain, singular)
            x = torch.where(better[:, None], candidate, x)
            normal = torch.where(better[:, None], trial, normal)
            cost = torch.where(better, trial_cost, cost)
            _, _, gradient, _ = solve_damped(normal, operator.scale(x), None, columns, step=False)
            history.append(cost.sum())
            if retiring:
                staying = ~(short & ~singular)
                x, live, data, normal, cost, damping, gradient = self._retire(staying, answer, live, x, data, normal,
                                                                              cost, damping, gradient)
                if x.shape[0] == 0:
                    break
            elif bool((short & ~singular).all()):
                break
        unconverged = int(x.shape[0]) if retiring else 0
        if retiring:
            if unconverged:
                answer[live] = x
        else:
            answer = x
        return Solution(
            x=answer.reshape(*shape, columns),
            cost=torch.stack(history),
            damping=torch.stack(weights) if weights else torch.zeros(0, device=start.device),
            iterations=iterations,
            unconverged=unconverged,
        )

    def _retire(self, keep, answer, live, x, data, normal, cost, damping, gradient):
        """Freeze the voxels that leave (their answer is their x) and compact
        the rest. One wait on the device, for how many stay; everything else
        is index arithmetic queued behind it."""
        kept = keep.nonzero().squeeze(1)
        if kept.numel() == keep.numel():
            return x, live, data, normal, cost, damping, gradient
        answer.index_copy_(0, live, torch.where(keep[:, None], answer.index_select(0, live), x))
        if hasattr(self.damping, "retire"):
            self.damping.retire(kept)

        def pick(t):
            return t.index_select(0, kept)

        return pick(x), pick(live), data, pick(normal), pick(cost), pick(damping), pick(gradient)
