"""Sequence design over the fused EPG kernels (reference optim/_design.py:85-287).

``crlb`` is the definition of the Cramer-Rao bound (diagonal of the inverse
Fisher matrix, real and imaginary channels as separate measurements), so it
reads as the reference's does; its gradient reaches the sequence through the
checkpointed adjoint kernel behind ``Simulator.jacobian``.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass
from typing import Any, Literal

import torch


@dataclass(frozen=True)
class Bounded:
    """An initial value and the limits it moves in (enforced by a sigmoid)."""

    initial: Any
    lower: Any
    upper: Any


@dataclass(frozen=True)
class SequenceOptimization:
    """The designed parameters and the loss at every step."""

    parameters: dict[str, torch.Tensor]
    loss: torch.Tensor


def crlb(
    jacobian: torch.Tensor,
    *,
    noise_variance: float = 1.0,
    singular: Literal["raise", "infinite"] = "raise",
) -> torch.Tensor:
    """Diagonal of the inverse Fisher matrix of ``(..., parameters, samples)``.

    Raises ``ValueError`` without a parameter axis; ``torch.linalg.LinAlgError``
    on a singular Fisher matrix unless ``singular="infinite"``.
    """
    if jacobian.ndim < 2:
        raise ValueError("a Jacobian is (..., parameters, samples)")
    if torch.is_complex(jacobian):
        rows = torch.cat((jacobian.real, jacobian.imag), dim=-1)
    else:
        rows = jacobian
    fisher = rows @ rows.transpose(-1, -2)
    if singular == "raise":
        inverse = torch.linalg.inv(fisher)
        return noise_variance * torch.diagonal(inverse, dim1=-2, dim2=-1)
    if singular != "infinite":
        raise ValueError(f"singular is 'raise' or 'infinite', got {singular!r}")
    inverse, failed = torch.linalg.inv_ex(fisher)
    bound = noise_variance * torch.diagonal(inverse, dim1=-2, dim2=-1)
    return bound.masked_fill((failed != 0)[..., None].expand_as(bound), float("inf"))


def _unconstrained(value: torch.Tensor, limits) -> torch.Tensor:
    if limits is None:
        return value.detach().clone()
    lower, upper = (torch.as_tensor(v, dtype=value.dtype, device=value.device) for v in limits)
    if bool(torch.any(lower >= upper)):
        raise ValueError("each lower limit must be below its upper limit")
    fraction = ((value - lower) / (upper - lower)).clamp(1e-6, 1 - 1e-6)
    return torch.logit(fraction).detach().clone()


def _constrained(raw: torch.Tensor, limits) -> torch.Tensor:
    if limits is None:
        return raw
    lower, upper = (torch.as_tensor(v, dtype=raw.dtype, device=raw.device) for v in limits)
    return lower + (upper - lower) * torch.sigmoid(raw)


class SequenceDesign(torch.nn.Module):
    """A cost and the bounded sequence parameters it is minimized over (Adam)."""

    def __init__(self, cost: Callable[..., torch.Tensor], **parameters: Any) -> None:
        super().__init__()
        if not parameters:
            raise ValueError("a design needs at least one parameter")
        self.cost = cost
        self._limits = {}
        raw = {}
        for name, entry in parameters.items():
            value = torch.as_tensor(entry.initial if isinstance(entry, Bounded) else entry)
            limits = (entry.lower, entry.upper) if isinstance(entry, Bounded) else None
            self._limits[name] = limits
            raw[name] = torch.nn.Parameter(_unconstrained(value, limits))
        self.raw = torch.nn.ParameterDict(raw)

    def values(self) -> dict[str, torch.Tensor]:
        return {name: _constrained(value, self._limits[name]) for name, value in self.raw.items()}

    def forward(self) -> torch.Tensor:
        loss = self.cost(**self.values())
        if loss.numel() != 1:
            raise ValueError("a design cost must return one number")
        return loss

    def minimize(self, *, iterations: int = 200, learning_rate: float = 0.05, optimizer_factory=None,
                 callback=None) -> SequenceOptimization:
        if iterations < 1 or learning_rate <= 0:
            raise ValueError("iterations and learning_rate must be positive")
        variables = list(self.raw.values())
        optimizer = torch.optim.Adam(variables, lr=learning_rate) if optimizer_factory is None else optimizer_factory(variables)
        history = []
        for step in range(iterations):
            optimizer.zero_grad(set_to_none=True)
            loss = self()
            loss.backward()
            optimizer.step()
            history.append(loss.detach())
            if callback is not None and callback(step, {k: v.detach() for k, v in self.values().items()}, loss.detach()):
                break
        return SequenceOptimization({k: v.detach() for k, v in self.values().items()}, torch.stack(history))
