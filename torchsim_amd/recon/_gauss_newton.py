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
    J^T r, J^T J in the free variables' natural columns; ``data`` holds the
    measured rows of the voxels ``live`` lists), ``scale(x)`` (the chain-rule
    factor d natural / d free) and ``columns``.
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
        if getattr(self.damping, "per_voxel", False):
            x, cost, weights, iterations, unconverged = self._retiring(operator, data, start)
        else:
            x, cost, weights, iterations, unconverged = self._together(operator, data, start)
        return Solution(
            x=x.reshape(*shape, start.shape[-1]),
            cost=cost,
            damping=weights,
            iterations=iterations,
            unconverged=unconverged,
        )

    def _step(self, operator, data, live, start, x, normal, cost, damping, columns):
        """One damped step of every voxel in ``x``: the candidate, its normal
        equations, the accepted state and the voxels whose step was short."""
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
        return candidate, trial, trial_cost, better, damping, short & ~singular

    def _together(self, operator, data, start):
        """Every voxel steps until all are flat or all steps are short."""
        columns = start.shape[-1]
        live = torch.arange(start.shape[0], device=start.device)
        x = start.clone()
        normal = operator.normal(x, data, live)
        cost = normal[:, 0].clone()
        _, _, gradient, curvature = solve_damped(normal, operator.scale(x), None, columns, step=False)
        damping = self.damping.begin(curvature)
        history, weights, iterations = [cost.sum()], [], 0
        for _ in range(self.max_iterations):
            if bool((gradient.abs().amax(-1) < self.gradient_tolerance).all()):
                break
            iterations += 1
            weights.append(damping.detach().reshape(-1).mean())
            candidate, trial, trial_cost, better, damping, short = self._step(
                operator, data, live, start, x, normal, cost, damping, columns)
            x = torch.where(better[:, None], candidate, x)
            normal = torch.where(better[:, None], trial, normal)
            cost = torch.where(better, trial_cost, cost)
            _, _, gradient, _ = solve_damped(normal, operator.scale(x), None, columns, step=False)
            history.append(cost.sum())
            if bool(short.all()):
                break
        stacked = torch.stack(weights) if weights else torch.zeros(0, device=start.device)
        return x, torch.stack(history), stacked, iterations, 0

    def _retiring(self, operator, data, start):
        """Each voxel leaves once its gradient is flat or its step short
        (reference recon/_gauss_newton.py:759-793), its answer frozen there.

        Leaving is a mask: a voxel that left keeps its x and is computed with
        the rest until the arrays are compacted, every COMPACT_EVERY
        iterations, so the loop waits on the device once per compaction rather
        than at every retirement. The history and the mean damping count the
        voxels still moving only, and the steps taken are those that began
        with a voxel still moving, as in a loop that compacts every time."""
        columns = start.shape[-1]
        device = start.device
        answer = start.clone()
        live = torch.arange(start.shape[0], device=device)
        x = start.clone()
        normal = operator.normal(x, data, live)
        cost = normal[:, 0].clone()
        _, _, gradient, curvature = solve_damped(normal, operator.scale(x), None, columns, step=False)
        damping = self.damping.begin(curvature)
        moving = torch.ones(start.shape[0], dtype=torch.bool, device=device)
        history, weights, began = [cost.sum()], [], []
        for iteration in range(self.max_iterations):
            moving = moving & ~(gradient.abs().amax(-1) < self.gradient_tolerance)
            if iteration % self.COMPACT_EVERY == 0:
                x, live, data, normal, cost, damping, gradient, moving = self._compact(
                    answer, live, x, data, normal, cost, damping, gradient, moving)
                if x.shape[0] == 0:
                    break
            count = moving.sum()
            began.append(count > 0)
            weights.append(torch.where(moving, damping.reshape(-1), 0.0).sum() / count.clamp_min(1))
            candidate, trial, trial_cost, better, damping, short = self._step(
                operator, data, live, start, x, normal, cost, damping, columns)
            accept = better & moving
            x = torch.where(accept[:, None], candidate, x)
            normal = torch.where(accept[:, None], trial, normal)
            cost = torch.where(accept, trial_cost, cost)
            _, _, gradient, _ = solve_damped(normal, operator.scale(x), None, columns, step=False)
            history.append(torch.where(moving, cost, 0.0).sum())
            moving = moving & ~short
        x, live, data, normal, cost, damping, gradient, moving = self._compact(
            answer, live, x, data, normal, cost, damping, gradient, moving)
        unconverged = int(x.shape[0])
        if unconverged:
            answer[live] = x
        ran = torch.stack(began).cpu() if began else torch.zeros(0, dtype=torch.bool)
        # a step that began with nobody moving was not taken; those come last
        taken = int(ran.sum())
        history = torch.stack(history[: 1 + taken])
        stacked = torch.stack(weights[:taken]) if taken else torch.zeros(0, device=device)
        return answer, history, stacked, taken, unconverged

    #: Iterations between compactions of the voxels still moving.
    COMPACT_EVERY = 4

    def _compact(self, answer, live, x, data, normal, cost, damping, gradient, moving):
        """Freeze the voxels that left (their answer is their x) and keep the
        rest, measured rows included: one wait on the device, for how many
        stay."""
        kept = moving.nonzero().squeeze(1)
        if kept.numel() == moving.numel():
            return x, live, data, normal, cost, damping, gradient, moving
        answer.index_copy_(0, live, torch.where(moving[:, None], answer.index_select(0, live), x))
        if hasattr(self.damping, "retire"):
            self.damping.retire(kept)

        def pick(t):
            return t.index_select(0, kept)

        return pick(x), pick(live), pick(data), pick(normal), pick(cost), pick(damping), pick(gradient), pick(moving)
