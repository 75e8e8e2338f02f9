"""Bounded parameters through a change of variables (reference _bounds.py:51-207).

A two-sided bound is a scaled sigmoid of the free variable, a one-sided one a
softplus; the derivative of the parameter with respect to its free variable
multiplies the Jacobian. The transforms are the reference's (so a fit lands on
the same iterates); they are the standard ones for the purpose.
"""

from __future__ import annotations

from collections.abc import Mapping, Sequence
from typing import Any

import torch

EDGE = 1e-6


def bound_of(bounds: Mapping[str, Any], name: str) -> tuple[float | None, float | None]:
    low, high = bounds.get(name, (None, None))
    if low is not None and high is not None and not high > low:
        raise ValueError(f"{name}: the bound ({low}, {high}) is not increasing")
    return low, high


def _softplus_inverse(y: torch.Tensor) -> torch.Tensor:
    return y + torch.log(-torch.expm1(-y))


def to_natural(free: torch.Tensor, bounds: Mapping[str, Any], names: Sequence[str]) -> torch.Tensor:
    if not bounds:
        return free
    columns = []
    for index, name in enumerate(names):
        low, high = bound_of(bounds, name)
        value = free[..., index]
        if low is not None and high is not None:
            columns.append(low + (high - low) * torch.sigmoid(value))
        elif low is not None:
            columns.append(low + torch.nn.functional.softplus(value))
        elif high is not None:
            columns.append(high - torch.nn.functional.softplus(-value))
        else:
            columns.append(value)
    return torch.stack(columns, dim=-1)


def to_free(natural: torch.Tensor, bounds: Mapping[str, Any], names: Sequence[str]) -> torch.Tensor:
    if not bounds:
        return natural
    columns = []
    for index, name in enumerate(names):
        low, high = bound_of(bounds, name)
        value = natural[..., index]
        if low is not None and high is not None:
            inside = ((value - low) / (high - low)).clamp(EDGE, 1.0 - EDGE)
            columns.append(torch.log(inside) - torch.log1p(-inside))
        elif low is not None:
            columns.append(_softplus_inverse((value - low).clamp_min(EDGE)))
        elif high is not None:
            columns.append(-_softplus_inverse((high - value).clamp_min(EDGE)))
        else:
            columns.append(value)
    return torch.stack(columns, dim=-1)


def widen(bounds: Mapping[str, Any], names: Sequence[str], free: torch.Tensor) -> torch.Tensor:
    """d(natural)/d(free), per column."""
    columns = []
    for index, name in enumerate(names):
        low, high = bound_of(bounds, name)
        value = free[..., index]
        if low is not None and high is not None:
            s = torch.sigmoid(value)
            columns.append((high - low) * s * (1.0 - s))
        elif low is not None:
            columns.append(torch.sigmoid(value))
        elif high is not None:
            columns.append(torch.sigmoid(-value))
        else:
            columns.append(torch.ones_like(value))
    return torch.stack(columns, dim=-1)
