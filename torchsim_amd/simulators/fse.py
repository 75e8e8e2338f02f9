"""Fast spin echo (reference simulators/fse.py:17-133)."""

from __future__ import annotations

from collections.abc import Mapping
from typing import Any

import torch

from ..model._simulator import Simulator, SpinPhysics
from ..sequence import _engine
from ..sequence._events import REFOCUSED, host_values
from .mrf import _matched


class FSESimulator(Simulator):
    """A refocused echo train sampled at every echo, then recovery over the
    rest of TR: ``M0 s (1 - E) / (1 - E s)`` with ``E = exp(-(TR - ESP n)/T1)``.

    The train runs on the complex kernel (the excitation and refocusing
    pulses sit a quarter turn apart); the recovery is applied to the samples,
    and its derivative joins the kernel's Jacobian by the chain rule.
    """

    model = SpinPhysics(properties={"T1": "t1_ms", "T2": "t2_ms", "M0": None, "B1": "b1"}, operators=REFOCUSED)
    states = 10
    signal_factor = None

    def layout(self, *, flip: Any, ESP: Any, phases: Any = 0.0, exc_flip: float = 90.0, exc_phase: float = 90.0,
               TR: Any = 1e6) -> list:
        angles = torch.deg2rad(torch.atleast_1d(torch.as_tensor(flip)))
        turns = torch.deg2rad(_matched(phases, angles))
        spacing_s = ESP * 1e-3
        angles, turns = host_values(angles, turns)
        ops = self.model.operators
        parts = [(0.0, ops.excitation(torch.pi / 180.0 * exc_flip, torch.pi / 180.0 * exc_phase))]
        for index in range(angles.shape[-1]):
            echo_s = (index + 1) * spacing_s
            parts.append((echo_s - 0.5 * spacing_s, ops.refocusing(angles[..., index], turns[..., index])))
            parts.append((echo_s, ops.readout(turns[..., index])))
        return parts

    def _recovery(self, properties: Mapping[str, Any], sequence: Mapping[str, Any], device) -> tuple:
        played = self.played(**sequence)
        echoes = torch.atleast_1d(torch.as_tensor(played["flip"])).shape[-1]
        left_ms = torch.as_tensor(played.get("TR", 1e6), dtype=torch.float32) - torch.as_tensor(
            played["ESP"], dtype=torch.float32) * echoes
        t1 = _engine._on_device(properties["T1"], device).reshape(-1, 1)
        if left_ms.dim() != 0 or left_ms.requires_grad:
            left_ms = left_ms.to(device)
        # (a 0-dim host tensor enters device arithmetic as a scalar operand)
        recovered = torch.exp(-1e3 / t1 * left_ms * 1e-3)
        density = _engine._on_device(properties.get("M0", 1.0), device).reshape(-1, 1)
        return recovered, density, left_ms, t1

    def evaluate(self, properties: Mapping[str, Any], **sequence: Any) -> torch.Tensor:
        signal, _ = self._run(properties, sequence)
        recovered, density, _, _ = self._recovery(properties, sequence, signal.device)
        return density * signal * (1 - recovered) / (1 - recovered * signal)

    def linearize(self, held, names, **sequence):
        tissue_names = tuple(name for name in names if name != "M0")
        signal, raw = self._run(held, sequence, tissue_names)
        recovered, density, left_ms, t1 = self._recovery(held, sequence, signal.device)
        denominator = 1 - recovered * signal
        out = density * signal * (1 - recovered) / denominator
        d_signal = density * (1 - recovered) / denominator**2
        d_recovered = density * signal * (signal - 1) / denominator**2
        columns = []
        for name in names:
            if name == "M0":
                columns.append(signal * (1 - recovered) / denominator)
                continue
            column = d_signal * raw[:, tissue_names.index(name)]
            if name == "T1":
                column = column + d_recovered * recovered * left_ms / (t1 * t1)
            columns.append(column)
        return out, torch.stack(columns, dim=1)
