"""Running one model from a flat argument list (reference _functional/_run.py:15-49)."""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any

import torch

from ..model._simulator import Simulator


def evaluated(
    model: Simulator,
    diff: str | Sequence[str] | None,
    device: str | torch.device | None = None,
    **values: Any,
) -> Any:
    """The signal, or the signal and its Jacobian when ``diff`` names properties."""
    given = {name: value for name, value in values.items() if value is not None}
    if device is not None:
        given["device"] = device
    if diff is None:
        return model.simulate(**given)
    return model.jacobian(diff, **given)
