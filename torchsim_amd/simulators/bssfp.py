"""Balanced SSFP in closed form (reference simulators/bssfp.py:18-121).

This model runs no EPG: a balanced repetition returns the transverse
magnetization to where it started, so the steady state is a function of the
phase a voxel accrues per TR and is evaluated elementwise over voxels and
contrasts. The EPG balanced train (a transient from equilibrium) is
:class:`TrueFISPSimulator`.
"""

from __future__ import annotations

import math
from collections.abc import Mapping, Sequence
from typing import Any

import torch

from ..model._simulator import Simulator, SpinPhysics, _as_torch, _backend, _like, _target_device

_SEQUENCE = ("flip", "TR", "TE", "phase_inc")


def steady_state(t1: torch.Tensor, t2: torch.Tensor, m0: torch.Tensor, b0: torch.Tensor, chemshift: torch.Tensor,
                 flip: torch.Tensor, tr: torch.Tensor, te: torch.Tensor, phase_inc: torch.Tensor) -> torch.Tensor:
    """Complex transverse magnetization a time TE after the pulse, all inputs
    broadcast together (T1, T2, TR, TE in ms; B0, chemshift in Hz; flip and
    phase_inc in degrees).

    With w = 2 pi (chemshift - B0), accrued phase theta = w TR + phase_inc,
    E1 = exp(-TR/T1) and e = exp(-TE/T2) (the reference takes the transverse
    decay at TE both inside the steady state and for the echo,
    simulators/bssfp.py:98-120):

        s = i M0 (1 - E1) sin(flip) (1 - e exp(-i theta)) / D * e * exp(i w TE),
        D = (1 - E1 cos flip)(1 - e cos theta) - e (E1 - cos flip)(e - cos theta).
    """
    alpha = torch.deg2rad(flip)
    omega = (2.0 * math.pi) * (chemshift - b0)
    theta = omega * (tr * 1e-3) + torch.deg2rad(phase_inc)
    e1 = torch.exp(-tr / t1)
    e = torch.exp(-te / t2)
    cos_a, cos_t = torch.cos(alpha), torch.cos(theta)
    denominator = (1.0 - e1 * cos_a) * (1.0 - e * cos_t) - e * (e1 - cos_a) * (e - cos_t)
    amplitude = m0 * (1.0 - e1) * torch.sin(alpha) * e / denominator
    # i (1 - e exp(-i theta)) = -e sin(theta) + i (1 - e cos(theta))
    shape = torch.complex(-e * torch.sin(theta), 1.0 - e * cos_t)
    echo = omega * (te * 1e-3)
    return amplitude * shape * torch.complex(torch.cos(echo), torch.sin(echo))


class bSSFPSimulator(Simulator):  # noqa: N801 - the reference's public name
    """The balanced SSFP steady state read at TE (TE defaults to TR/2).

    Properties: T1, T2 (ms), M0, B0 and chemshift (Hz). Sequence: flip (deg),
    TR, TE (ms), phase_inc (deg, linear phase cycling, default 180). Every
    sequence argument is a scalar or one value per contrast; the signal is
    ``(..., contrasts)``, or one value per voxel when every sequence argument
    is a scalar. Evaluated with torch elementwise arithmetic on the device
    the call names (a GPU, as every model of this library).
    """

    model = SpinPhysics(properties={"T1": None, "T2": None, "M0": None, "B0": None, "chemshift": None})
    signal_factor = None

    def _evaluate(self, held: Mapping[str, Any], sequence: Mapping[str, Any],
                  tangent: str | None = None) -> tuple[torch.Tensor, torch.Tensor | None]:
        if "T1" not in held or "T2" not in held:
            raise ValueError("bSSFPSimulator needs T1 and T2")
        unknown = sorted(set(sequence) - set(_SEQUENCE) - {"device"})
        if unknown:
            raise TypeError(f"bSSFPSimulator takes no sequence arguments {unknown}")
        device = _target_device({**held, **sequence}, sequence.get("device"))
        cast = lambda v: torch.as_tensor(v, dtype=torch.float32).to(device)  # noqa: E731
        tissue = {name: cast(held.get(name, default)) for name, default in
                  (("T1", None), ("T2", None), ("M0", 1.0), ("B0", 0.0), ("chemshift", 0.0))}
        tr = cast(sequence["TR"])
        te = tr / 2 if sequence.get("TE") is None else cast(sequence["TE"])
        played = {"flip": cast(sequence["flip"]), "tr": tr, "te": te,
                  "phase_inc": cast(sequence.get("phase_inc", 180.0))}
        contrasts = torch.broadcast_shapes(*(value.shape for value in played.values()))
        if len(contrasts) > 1:
            raise ValueError("sequence arguments are scalars or one value per contrast")
        if contrasts:  # the contrast axis trails the voxel axes
            tissue = {name: value[..., None] for name, value in tissue.items()}

        def signal_of(t1, t2, m0, b0, chemshift):
            return steady_state(t1, t2, m0, b0, chemshift, played["flip"], played["tr"], played["te"],
                                played["phase_inc"])

        order = ("T1", "T2", "M0", "B0", "chemshift")
        primals = tuple(tissue[name] for name in order)
        if tangent is None:
            return signal_of(*primals), None
        tangents = tuple(torch.ones_like(value) if name == tangent else torch.zeros_like(value)
                         for name, value in zip(order, primals))
        return torch.func.jvp(signal_of, primals, tangents)

    def simulate(self, **values: Any) -> Any:
        backend = _backend(values)
        held, sequence = self._split({**self.bound, **values})
        signal, _ = self._evaluate(held, sequence)
        return _like(signal, backend)

    def jacobian(self, diff: str | Sequence[str], **values: Any) -> tuple[Any, Any]:
        """The signal and its derivative along each name in ``diff`` (forward
        mode, one pass per name); a single name drops the name axis.

        Raises
        ------
        ValueError
            If ``diff`` names a property the call did not give.
        """
        names = (diff,) if isinstance(diff, str) else tuple(diff)
        backend = _backend(values)
        held, sequence = self._split({**self.bound, **values})
        for name in names:
            if name not in held:
                raise ValueError(f"{name!r} is not among the properties this call gave "
                                 f"bSSFPSimulator: {sorted(held)}")
        columns = []
        signal = None
        for name in names:
            signal, column = self._evaluate(held, sequence, tangent=name)
            columns.append(column)
        if signal is None:
            signal, _ = self._evaluate(held, sequence)
        shape = torch.broadcast_shapes(signal.shape, *(c.shape for c in columns))
        # names sit before the contrast axis, or last when there is none
        axis = -2 if self._has_contrasts(sequence) else -1
        stacked = torch.stack([c.expand(shape) for c in columns], dim=axis)
        if isinstance(diff, str):
            stacked = stacked.squeeze(axis)
        return _like(signal.expand(shape), backend), _like(stacked, backend)

    @staticmethod
    def _has_contrasts(sequence: Mapping[str, Any]) -> bool:
        return any(torch.as_tensor(_as_torch(sequence[name])).dim() > 0 for name in _SEQUENCE
                   if sequence.get(name) is not None)
