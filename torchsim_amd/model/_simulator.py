"""Signal models: declared properties, a layout of operators, and the two
derivative modes over them.

Mirrors the reference's public model API (model/_signal.py:158-240
``simulate``/``jacobian``; model/_state_machine.py:356-916 ``Simulator``):
properties and sequence arguments are given together and told apart by the
model's declared properties; a scalar-only call returns one signal, a call with
maps returns one per voxel; ``jacobian(diff)`` returns the signal and its
derivative along each named property (a single name collapses that axis).

What differs is below the API: a model whose evaluate is the recorded sample
times a constant (``signal_factor``) runs its signal and every Jacobian column
in ONE fused gfx950 launch, where the reference issues one forward-mode pass
per property.
"""

from __future__ import annotations

import threading
from collections import OrderedDict
from collections.abc import Mapping, Sequence
import dataclasses
from copy import copy as shallow_copy
from dataclasses import dataclass, field
from types import SimpleNamespace
from typing import Any

import numpy as np
import torch
from torch.autograd import forward_ad

from .. import _native
from ..sequence import _dephasing, _engine
from ..sequence._events import INVERSION, EventOperators, PackedEvents, compose, pack

#: Public property names a call may give beside a model's own (reference
#: sequence/_parameters.py PUBLIC_PROPERTIES, the subset this library carries).
PUBLIC_PROPERTIES = {
    "T1": "t1_ms",
    "T2": "t2_ms",
    "M0": "m0",
    "B1": "b1",
    "inv_efficiency": "inversion_efficiency",
    "D": "diffusion_um2_per_ms",
    "B0": "b0_hz",
    "B1phase": "b1_phase_rad",
    "poolB_fraction": "pool_b_fraction",
    "poolB_exchange": "pool_b_exchange_rate",
    "poolB_T1": "t1_pool_b_ms",
    "poolB_T2": "t2_pool_b_ms",
    "poolB_shift": "pool_b_shift_hz",
    # the same pool, also accepted under these names
    "pool_b_fraction": "pool_b_fraction",
    "pool_b_exchange": "pool_b_exchange_rate",
    "T1_pool_b": "t1_pool_b_ms",
    "T2_pool_b": "t2_pool_b_ms",
    "pool_b_shift": "pool_b_shift_hz",
    # the semisolid (magnetization-transfer) pool (reference
    # sequence/_parameters.py:117-123, 245-247)
    "bound_fraction": "bound_fraction",
    "bound_exchange": "bound_exchange_rate",
    "bound_T1": "t1_bound_ms",
    # the field's spread across the voxel: reaches the samples, not the states
    # (reference sequence/_parameters.py:168-173, 242)
    "T2prime": "t2_prime_ms",
}

#: Properties the sample factor carries rather than the kernels, where the
#: stream's winding admits it (B0 otherwise stays in the states).
SAMPLE_FIELDS = ("b0_hz", "t2_prime_ms")

#: Run settings a call may name (reference model/_state_machine.py:97-106),
#: and the gradient geometry a diffusion coefficient is read through
#: (reference sequence/_accelerators.py:371-420).
RUN_SETTINGS = (
    "states", "repetitions", "device", "tile", "checkpoint_pulses", "checkpoint_events", "crusher_dephasing_rad",
    "voxel_size_m",
    # a shaped pulse the layout's pulses play, and where across the slice to
    # work it out (reference model/_state_machine.py:399, 500-513, 552-553)
    "pulse", "across_slice",
)

#: The apparent diffusion coefficient is given in um^2/ms; the b-factor wants m^2/s.
_DIFFUSION_UNIT = 1e-9


def damping_scale(sequence: Mapping[str, Any]) -> float:
    """What a diffusion coefficient is multiplied by to give the kernels' rate:
    (crusher dephasing / voxel size)^2 * 1e-9, zero without a declared crusher
    (reference sequence/_accelerators.py:374-398)."""
    crusher = sequence.get("crusher_dephasing_rad")
    if not crusher:
        return 0.0
    voxel = sequence.get("voxel_size_m")
    if voxel is None:
        raise ValueError(
            "a crusher winds its dephasing across a voxel, so a sequence that declares "
            "crusher_dephasing_rad must declare voxel_size_m too"
        )
    per_m = float(crusher) / float(voxel)
    return per_m * per_m * _DIFFUSION_UNIT


_PACKED_CACHE: OrderedDict = OrderedDict()
_PACKED_CACHE_SIZE = 256
_PACKED_MEMO_SIZE = 8
_PACKED_LOCK = threading.Lock()


def _protocol_key(kind: type, sequence: Mapping[str, Any]) -> tuple | None:
    """A hashable key for a protocol, or None if some argument carries a graph."""
    parts = [kind.__qualname__]
    for name in sorted(sequence):
        if name in ("device", "tile"):
            continue
        value = sequence[name]
        if isinstance(value, torch.Tensor):
            if _engine.carries_derivative(value):
                return None
            held = value.detach().cpu().contiguous()
            parts.append((name, str(held.dtype), tuple(held.shape), held.numpy().tobytes()))
        elif isinstance(value, np.ndarray):
            parts.append((name, str(value.dtype), value.shape, np.ascontiguousarray(value).tobytes()))
        elif isinstance(value, (int, float, complex, str, bool, type(None))):
            parts.append((name, type(value).__name__, value))
        elif hasattr(value, "key") and hasattr(value, "envelope"):
            parts.append((name, "pulse", value.key))  # a shaped pulse, by its samples
        elif dataclasses.is_dataclass(value) and getattr(type(value), "__hash__", None) is not None:
            parts.append((name, type(value).__name__, repr(value)))
        else:
            return None
    return tuple(parts)


@dataclass(frozen=True)
class SpinPhysics:
    """Which properties a voxel exposes and what each kind of event plays."""

    properties: Mapping[str, str | None] = field(default_factory=dict)
    operators: EventOperators = field(default_factory=EventOperators)


def _as_torch(value: Any) -> Any:
    if isinstance(value, torch.Tensor):
        return value
    if isinstance(value, np.ndarray):
        return torch.from_numpy(value)
    if isinstance(value, (list, tuple)):
        return torch.as_tensor(value)
    return value


def _backend(values: Mapping[str, Any]) -> Any:
    """numpy if the first array the caller brought is numpy, else None."""
    for value in values.values():
        if isinstance(value, np.ndarray):
            return np
        if isinstance(value, torch.Tensor):
            return None
    return None


def _like(result: torch.Tensor, backend: Any) -> Any:
    if backend is np:
        return result.detach().cpu().numpy()
    return result


def _target_device(values: Mapping[str, Any], device: Any) -> torch.device:
    if device is not None:
        device = torch.device(device)
        if device.type == "cuda" and device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        return device
    for value in values.values():
        if isinstance(value, torch.Tensor) and value.device.type == "cuda":
            return value.device
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU is visible: torchsim_amd has no CPU path")
    return torch.device("cuda", torch.cuda.current_device())


class Simulator:
    """A signal model made of an event layout run through the EPG engine.

    Subclasses declare ``model`` (a :class:`SpinPhysics`), ``states`` (the
    default number of configuration orders) and ``layout(**sequence)``; they
    may override ``evaluate`` to post-process the recorded samples. A model
    whose recorded signal is the sample times a constant declares that
    constant as ``signal_factor`` and gets the fused Jacobian.
    """

    model: SpinPhysics = SpinPhysics()
    states: int | None = None
    repetitions: int = 1
    signal_factor: complex | None = 1.0

    def __init__(self, **values: Any) -> None:
        self.bound = {name: _as_torch(value) if name in self.exposes else value for name, value in values.items()}

    # -- what is fixed on the model -------------------------------------------

    @property
    def exposes(self) -> tuple[str, ...]:
        return tuple(self.model.properties)

    @property
    def accepts(self) -> tuple[str, ...]:
        return tuple(dict.fromkeys((*self.exposes, *PUBLIC_PROPERTIES)))

    def bind(self, **values: Any) -> Simulator:
        held = shallow_copy(self)
        held.bound = {**self.bound, **{name: _as_torch(value) if name in self.exposes else value for name, value in values.items()}}
        return held

    def to(self, device: torch.device | str) -> Simulator:
        moved = shallow_copy(self)
        moved.bound = {name: value.to(device) if isinstance(value, torch.Tensor) else value for name, value in self.bound.items()}
        return moved

    # -- layout ------------------------------------------------------------------

    def layout(self, **sequence: Any) -> list:  # pragma: no cover - abstract
        raise NotImplementedError

    def played(self, **sequence: Any) -> dict[str, Any]:
        """The sequence arguments the layout takes (run settings removed)."""
        return {name: value for name, value in sequence.items() if name not in RUN_SETTINGS}

    def repetition_us(self, played_us: Any, **sequence: Any) -> Any:
        """How long one playing of the train lasts, for repeated playings."""
        return played_us

    def packed(self, **sequence: Any) -> PackedEvents:
        """The packed event stream of this protocol (host tensors).

        A protocol whose arguments carry no autograd graph is resolved once and
        remembered (the reference resolves a structure once per protocol too,
        model/_binding.py:94-139); later calls with equal arguments reuse it.
        """
        key = _protocol_key(type(self), sequence)
        if key is None:
            return self._packed(**sequence)
        # the protocols this simulator has played stay with it (a request in
        # flight never loses its stream to the shared cache's eviction)
        memo = self.__dict__.setdefault("_packed_memo", {})
        cached = memo.get(key)
        if cached is not None:
            return cached
        with _PACKED_LOCK:
            cached = _PACKED_CACHE.get(key)
            if cached is not None:
                _PACKED_CACHE.move_to_end(key)
        if cached is None:
            cached = self._packed(**sequence)
            with _PACKED_LOCK:
                _PACKED_CACHE[key] = cached
                while len(_PACKED_CACHE) > _PACKED_CACHE_SIZE:
                    _PACKED_CACHE.popitem(last=False)
        if len(memo) >= _PACKED_MEMO_SIZE:
            memo.clear()
        memo[key] = cached
        return cached

    def _packed(self, **sequence: Any) -> PackedEvents:
        played = self.played(**sequence)
        events, end_s = compose(*self.layout(**played))
        repetitions = int(sequence.get("repetitions", self.repetitions) or 1)
        tr_us = self.repetition_us(1e6 * end_s, **played) if repetitions > 1 else 0.0
        packed = pack(events, repetitions=repetitions, tr_duration_us=tr_us)
        pulse = sequence.get("pulse")
        if pulse is not None:
            # a shaped pulse deposits its own power in a semisolid pool, and
            # its area's phase turns every pulse it plays (reference
            # sequence/_accelerators.py:1016-1025, 1113-1125)
            rf = packed.kind == 1
            turned = rf & ((packed.action & INVERSION) == 0)
            packed.saturation = torch.where(rf, torch.full_like(packed.saturation, pulse.saturation()),
                                            packed.saturation)
            if pulse.integral_phase() != 0.0:
                packed.phase = packed.phase + turned.to(packed.phase.dtype) * pulse.integral_phase()
        return packed

    # -- the two derivative modes -----------------------------------------------

    def _split(self, values: Mapping[str, Any]) -> tuple[dict, dict]:
        declared = self.accepts
        held = {name: _as_torch(values[name]) for name in declared if name in values}
        sequence = {name: value for name, value in values.items() if name not in declared}
        return held, sequence

    def _tissue_fields(self, held: Mapping[str, Any]) -> dict[str, Any]:
        fields: dict[str, Any] = {}
        for name, value in held.items():
            target = self.model.properties.get(name, PUBLIC_PROPERTIES.get(name)) if name in self.model.properties else PUBLIC_PROPERTIES.get(name)
            if target is not None:
                fields[target] = value
        return fields

    @staticmethod
    def _batch(held: Mapping[str, Any]) -> tuple[int, ...]:
        shapes = [torch.as_tensor(value).shape for value in held.values()]
        if not shapes:
            return ()
        return tuple(torch.broadcast_shapes(*shapes))

    def _prepared(self, held: Mapping[str, Any], sequence: Mapping[str, Any], diff: tuple[str, ...] = (),
                  device: Any = None) -> tuple:
        """The host side of a run: the device, the packed protocol, the tissue
        fields and their prepared buffers, the state count, which fields the
        kernels differentiate (forward-mode columns), which require grad or
        carry a forward-mode tangent, and the sample factor (a
        :class:`~torchsim_amd.sequence._dephasing.Dephasing`, or None) that
        carries B0 and T2' where the stream's winding admits it -- B0 is then
        held out of the kernels' tissue. ``device`` overrides the one the
        values name."""
        device = _target_device({**held, **sequence}, device if device is not None else sequence.get("device"))
        packed = self.packed(**sequence)
        fields = self._tissue_fields(held)
        if "diffusion_um2_per_ms" in fields:
            scale = damping_scale(sequence)
            value = fields["diffusion_um2_per_ms"]
            fields["diffusion_um2_per_ms"] = value * scale if scale else 0.0
        differentiated = tuple(self._tissue_fields({name: held[name] for name in diff}).keys())
        dephasing = self._dephasing(held, packed, fields, differentiated, device)
        if dephasing is not None:
            differentiated = tuple(name for name in differentiated if name not in SAMPLE_FIELDS)
        # fields autograd or a forward-mode tangent must reach are laid out per
        # atom, like forward-mode ones
        reversed_fields = tuple(name for name, value in fields.items()
                                if isinstance(value, torch.Tensor) and value.requires_grad)
        tangent_fields = tuple(name for name, value in fields.items()
                               if _engine.carries_derivative(value) and name not in reversed_fields)
        tissue = _engine.prepare_tissue(fields, device,
                                        tuple(dict.fromkeys(differentiated + reversed_fields + tangent_fields)))
        states = int(sequence.get("states", self.states) or 10)
        return device, packed, fields, tissue, states, differentiated, reversed_fields, tangent_fields, dephasing

    def _dephasing(self, held: Mapping[str, Any], packed: PackedEvents, fields: dict, differentiated: tuple,
                   device: torch.device) -> _dephasing.Dephasing | None:
        """Take B0 and T2' out of ``fields`` into a sample factor where they
        act and the winding admits it (reference
        sequence/_accelerators.py:647-652); T2' where it does not is refused.
        A property acts where it carries a derivative, is differentiated, or
        (B0) is nonzero / (T2') finite anywhere."""
        t2_prime = fields.pop("t2_prime_ms", None)
        b0 = fields.get("b0_hz")
        spreading = t2_prime is not None and (
            "t2_prime_ms" in differentiated or _engine.carries_derivative(t2_prime)
            or bool(torch.isfinite(torch.as_tensor(t2_prime)).any()))
        turning = b0 is not None and ("b0_hz" in differentiated or _engine.carries_derivative(b0)
                                      or _dephasing.applies(b0))
        if not spreading and not turning:
            return None
        if not _dephasing.winding(packed).analytic:
            if spreading:
                _dephasing.refuse_a_spread()
            return None
        fields.pop("b0_hz", None)
        shape = self._batch(held)

        def per_atom(value: Any) -> torch.Tensor:
            held_value = _engine._on_device(value, device).to(torch.float32)
            return held_value.expand(shape).reshape(-1)

        return _dephasing.Dephasing(b0_hz=per_atom(b0) if turning else None,
                                    t2_prime_ms=per_atom(t2_prime) if spreading else None,
                                    tau_s=_dephasing.unrefocused_s(packed, device))

    def _run(
        self,
        held: Mapping[str, Any],
        sequence: Mapping[str, Any],
        diff: tuple[str, ...] = (),
        subspace: Any = None,
    ) -> tuple[torch.Tensor, torch.Tensor | None]:
        """Signal (atoms, N) and Jacobian (atoms, P, N) of the recorded samples;
        with a ``subspace``, their projections (atoms, R) and (atoms, P, R)."""
        prepared = self._prepared(held, sequence, diff)
        dephasing = prepared[-1]
        if dephasing is None:
            return self._run_states(prepared, held, sequence, diff, subspace)
        if subspace is not None:
            # the factor differs per atom and sample: project what it made
            signal, jacobian = self._run(held, sequence, diff)
            return subspace.project(signal), (None if jacobian is None else subspace.project(jacobian))
        signal, jacobian = self._run_states(prepared, held, sequence, diff, None)
        factor = dephasing.factor()
        signal = signal * factor
        names = tuple(self._tissue_fields({name: held[name] for name in diff}).keys())
        if not names:
            return signal, None
        kernel_columns = prepared[5]
        columns = [dephasing.column(signal, name) if name in SAMPLE_FIELDS
                   else jacobian[:, kernel_columns.index(name)] * factor for name in names]
        return signal, torch.stack(columns, dim=1)

    @staticmethod
    def _profile(sequence: Mapping[str, Any]) -> Any:
        """The transition table of the run's shaped pulse, or None where its
        pulses turn as flips and phases name: a pulse that is slice-selective
        or moves its phase asks for one by itself (at the slice centre unless
        ``across_slice`` says where), and naming positions asks for one for
        any pulse (reference sequence/_accelerators.py:292-318)."""
        from ..sequence import _transition

        pulse = sequence.get("pulse")
        asked = _transition.across_the_slice(sequence.get("across_slice"))
        if pulse is None:
            if asked is not None:
                raise ValueError("across_slice says where to work out a shaped pulse; give the pulse as pulse=")
            return None
        if asked is None and not pulse.selective():
            return None
        profile = asked or _transition.exact_slice_profile(1)
        return _transition.transition_table(pulse, profile.positions(), bins=profile.bins,
                                            theta_max=profile.theta_max)

    def _run_states(self, prepared: tuple, held: Mapping[str, Any], sequence: Mapping[str, Any],
                    diff: tuple[str, ...], subspace: Any) -> tuple[torch.Tensor, torch.Tensor | None]:
        """The kernels' part of :meth:`_run`, on its prepared inputs."""
        device, packed, fields, tissue, states, differentiated, reversed_fields, tangent_fields, _ = prepared
        profile = self._profile(sequence)
        real, flip, samples = _engine.real_subspace(packed, tissue, device)
        real = real and profile is None
        factor = self.signal_factor if self.signal_factor is not None else 1.0
        sequence_live = [_engine.carries_derivative(v) for v in (packed.flip, packed.phase, packed.duration)]
        reverse = bool(reversed_fields) or any(v.requires_grad for v in (packed.flip, packed.phase, packed.duration))
        forward_mode = bool(tangent_fields) or (any(sequence_live) and not reverse)
        if subspace is not None:
            if subspace.contrasts != packed.output_count:
                raise ValueError(f"the subspace spans {subspace.contrasts} contrasts, the sequence records "
                                 f"{packed.output_count}")
            if real and not reverse and not forward_mode and _engine.projects_fused(tissue, differentiated,
                                                                                     subspace.rank):
                return _engine.simulate_projected(packed, tissue, subspace.basis, states=states,
                                                  columns=differentiated, factor=factor, flip=flip, samples=samples)
            signal, jacobian = self._run(held, sequence, diff)
            return subspace.project(signal), (None if jacobian is None else subspace.project(jacobian))
        from ..sequence import _complex

        if profile is not None or tissue.semisolid:
            from ..sequence import _semisolid

            if any(sequence_live):
                if differentiated:
                    raise NotImplementedError(
                        "a Jacobian through a semisolid pool or a profiled pulse is not differentiated again"
                    )
                # derivatives to the flips, phases and durations: the adjoint (reverse mode) and the
                # event tangents (forward mode) of this kernel
                return _semisolid.sequence_signal(
                    packed, tissue, profile, states=states,
                    carried=tuple(dict.fromkeys(reversed_fields + tangent_fields)), factor=factor), None
            carried = tuple(dict.fromkeys(reversed_fields + tangent_fields))
            if carried:
                if differentiated:
                    raise NotImplementedError(
                        "a Jacobian through a semisolid pool or a profiled pulse is not differentiated again"
                    )
                return _semisolid.carried_signal(packed, tissue, profile, states=states, carried=carried,
                                                 factor=factor), None
            return _semisolid.simulate_shaped(packed, tissue, profile, states=states, columns=differentiated,
                                              factor=factor)
        if not differentiated and (reverse or forward_mode):
            # the plain signal under reverse or forward mode: one Function whose
            # backward takes the adjoint that carries what requires grad and
            # whose jvp follows any direction
            carried = set(_complex.GRADIENT_FIELDS)
            unsupported = [name for name in reversed_fields + tangent_fields if name not in carried]
            if unsupported:
                raise NotImplementedError(
                    f"derivatives reach {', '.join(_complex.GRADIENT_FIELDS)} and the sequence's flips, phases "
                    f"and durations; {unsupported} also carry one"
                )
            signs = packed.__dict__.get("_flip_signs") if real else None
            run = _complex.SignalRun(
                packed=packed, tissue=tissue, states=states, factor=factor, real=real, flip_signs=signs,
                samples=samples, checkpoint_pulses=int(sequence.get("checkpoint_pulses", 64)),
                checkpoint_events=int(sequence.get("checkpoint_events", 0)),
            )
            signal = _complex.EpgSignal.apply(
                packed.flip.to(device), packed.phase.to(device), packed.duration.to(device), *tissue.buffers, run,
            )
            return signal, None
        if differentiated and forward_mode:
            raise NotImplementedError(
                "a Jacobian is already a set of forward-mode columns; take torch.func.jvp of simulate() instead"
            )
        if not real:
            if reverse:
                if reversed_fields:
                    raise NotImplementedError(
                        "a Jacobian under reverse mode on the complex kernel passes gradients to the sequence's "
                        f"flips, phases and durations; tissue maps are constants there ({list(reversed_fields)} "
                        "require grad)"
                    )
                run = SimpleNamespace(packed=packed, tissue=tissue, states=states, factor=factor,
                                      checkpoint_events=int(sequence.get("checkpoint_events", 0)))
                signal, jacobian = _complex.ComplexJacobian.apply(
                    packed.flip.to(device), packed.phase.to(device), packed.duration.to(device), run, differentiated)
                return signal, (jacobian if differentiated else None)
            return _complex.simulate_complex(packed, tissue, states=states, columns=differentiated, factor=factor)
        if reverse:
            unsupported = [name for name in reversed_fields if name not in _engine.GRADIENT_FIELDS]
            if unsupported or packed.phase.requires_grad:
                raise NotImplementedError(
                    f"a Jacobian under reverse mode reaches T1, T2, M0, B1, the inversion efficiency and the "
                    f"sequence's flips and durations; {unsupported or ['phase']} also require grad"
                )
            signs = packed.__dict__.get("_flip_signs")
            flip_in = packed.flip if signs is None else packed.flip * signs
            buffers = [tissue.buffers[_native.TISSUE_INDEX[name]] if name in reversed_fields else None
                       for name in _engine.GRADIENT_FIELDS]
            signal, jacobian = _engine.FusedRealJacobian.apply(
                flip_in.to(device), packed.duration.to(device), *buffers, packed, tissue, states, differentiated,
                factor, samples, int(sequence.get("checkpoint_pulses", 64)),
                _engine.zero_intervals_reached(packed) if packed.duration.requires_grad else None,
            )
            return signal, (jacobian if differentiated else None)
        return _engine.simulate_real(
            packed,
            tissue,
            states=states,
            columns=differentiated,
            factor=factor,
            flip=flip,
            samples=samples,
            tile=sequence.get("tile", "auto"),
        )

    def evaluate(self, properties: Mapping[str, Any], **sequence: Any) -> torch.Tensor:
        """The recorded signal, (voxels, samples)."""
        signal, _ = self._run(properties, sequence)
        return signal

    def simulate(self, **values: Any) -> Any:
        """Return the recorded signal, shaped like the caller's property maps."""
        if _transformed():
            # tensors bound at construction (a protocol's flips, TRs, phases) are
            # the Function's inputs too, so that a transform's tangents and
            # cotangents along them reach the kernels
            given = {**{name: value for name, value in self.bound.items() if isinstance(value, torch.Tensor)},
                     **values}
            keys = tuple(name for name, value in given.items() if isinstance(value, torch.Tensor))
            others = {name: value for name, value in given.items() if name not in keys}
            return _TransformBridge.apply(self, keys, others, *(given[name] for name in keys))
        backend = _backend(values)
        held, sequence = self._split({**self.bound, **values})
        batch = self._batch(held)
        signal = self.evaluate(held, **sequence)
        return _like(signal.reshape(*batch, signal.shape[-1]), backend)

    def jacobian(self, diff: str | Sequence[str], **values: Any) -> tuple[Any, Any]:
        """Return the signal and its derivative along each name in ``diff``.

        Raises
        ------
        ValueError
            If ``diff`` names a property the call did not give, or one that is
            not floating point.
        """
        names = (diff,) if isinstance(diff, str) else tuple(diff)
        if _transformed():
            raise NotImplementedError(
                "a Jacobian is already a set of forward-mode columns; under torch.func take the transform of "
                "simulate() instead"
            )
        backend = _backend(values)
        held, sequence = self._split({**self.bound, **values})
        for name in names:
            if name not in held:
                raise ValueError(
                    f"{name!r} is not among the properties this call gave "
                    f"{type(self).__name__}: {sorted(held)}. Forward mode "
                    f"differentiates a property the tissue carries; for a "
                    f"sequence argument, build a cost and call backward()."
                )
            if not torch.is_floating_point(torch.as_tensor(held[name])):
                raise ValueError(f"{name!r} does not carry a derivative")
        batch = self._batch(held)
        signal, jacobian = self.linearize(held, names, **sequence)
        signal = signal.reshape(*batch, signal.shape[-1])
        jacobian = jacobian.reshape(*batch, len(names), jacobian.shape[-1])
        if isinstance(diff, str):
            jacobian = jacobian[..., 0, :]
        return _like(signal, backend), _like(jacobian, backend)

    def project(self, subspace: Any, diff: str | Sequence[str] | None = None, **values: Any) -> Any:
        """The recorded signal projected onto ``subspace`` (a
        :class:`~torchsim_amd.Subspace` over this sequence's samples), and
        with ``diff`` its Jacobian projected too: what
        ``subspace.project(simulate(...))`` and ``subspace.project(jacobian(...))``
        return (reference recon/_operator.py:446-473), shaped like the
        caller's maps with the sample axis replaced by the rank.

        A real-subspace sequence without diffusion contracts the samples with
        the basis inside the kernel (ranks up to 16), so the full-length
        signal and Jacobian are never stored; any other runs the full
        simulation and projects it on the device.
        """
        names = () if diff is None else ((diff,) if isinstance(diff, str) else tuple(diff))
        backend = _backend(values)
        held, sequence = self._split({**self.bound, **values})
        for name in names:
            if name not in held:
                raise ValueError(f"{name!r} is not among the properties this call gave "
                                 f"{type(self).__name__}: {sorted(held)}")
            if not torch.is_floating_point(torch.as_tensor(held[name])):
                raise ValueError(f"{name!r} does not carry a derivative")
        signal, jacobian = self.linearize_projected(held, names, subspace, **sequence)
        batch = self._batch(held)
        signal = signal.reshape(*batch, signal.shape[-1])
        if not names:
            return _like(signal, backend)
        jacobian = jacobian.reshape(*batch, len(names), jacobian.shape[-1])
        if isinstance(diff, str):
            jacobian = jacobian[..., 0, :]
        return _like(signal, backend), _like(jacobian, backend)

    def linearize_projected(self, held: Mapping[str, Any], names: tuple[str, ...], subspace: Any,
                            **sequence: Any) -> tuple[torch.Tensor, torch.Tensor | None]:
        """Signal (voxels, R) and Jacobian (voxels, P, R) projected onto
        ``subspace`` (fused into the kernel where the model allows it)."""
        if self.signal_factor is not None:
            return self._run(held, sequence, tuple(names), subspace=subspace)
        if subspace.contrasts != self.packed(**sequence).output_count:
            raise ValueError(f"the subspace spans {subspace.contrasts} contrasts, the sequence records "
                             f"{self.packed(**sequence).output_count}")
        if not names:
            return subspace.project(self.evaluate(held, **sequence)), None
        signal, jacobian = self.linearize(held, tuple(names), **sequence)
        return subspace.project(signal), subspace.project(jacobian)

    def linearize(self, held: Mapping[str, Any], names: tuple[str, ...], **sequence: Any) -> tuple[torch.Tensor, torch.Tensor]:
        """Signal (voxels, N) and Jacobian (voxels, P, N) along ``names``."""
        if self.signal_factor is None:
            raise NotImplementedError(f"{type(self).__name__} must define linearize")
        signal, jacobian = self._run(held, sequence, names)
        return signal, jacobian


def _transformed() -> bool:
    """Whether a torch.func transform (jvp, grad, vjp) is running: host work
    (packing a protocol, its real-subspace verdict) cannot read tensors there,
    so a simulation then runs as one Function of the caller's tensors."""
    try:
        return torch._C._functorch.peek_interpreter_stack() is not None
    except AttributeError:  # pragma: no cover - a torch without functorch
        return False


class _TransformBridge(torch.autograd.Function):
    """``Simulator.simulate`` as one Function of the caller's tensors, for
    torch.func transforms: its forward runs the ordinary simulation on plain
    tensors; its backward runs it again under autograd (the adjoint kernels);
    its jvp carries the caller's tangents onto the packed events and the
    prepared tissue buffers (a double backward through those cheap host-side
    maps) and runs the JVP kernel."""

    @staticmethod
    def forward(simulator, keys, others, *tensors):
        return simulator.simulate(**others, **dict(zip(keys, tensors)))

    @staticmethod
    def setup_context(ctx, inputs, output):
        simulator, keys, others, *tensors = inputs
        ctx.simulator, ctx.keys, ctx.others = simulator, keys, others
        ctx.save_for_backward(*tensors)
        ctx.save_for_forward(*tensors)

    @staticmethod
    def backward(ctx, grad):
        needs = tuple(bool(n) for n in ctx.needs_input_grad[3:])
        grads = _BridgeGradient.apply(ctx.simulator, ctx.keys, ctx.others, needs, grad, *ctx.saved_tensors)
        return (None, None, None, *grads)

    @staticmethod
    def jvp(ctx, _simulator, _keys, _others, *tangents):
        return _BridgeTangent.apply(ctx.simulator, ctx.keys, ctx.others, len(tangents), *ctx.saved_tensors,
                                    *tangents)


class _BridgeGradient(torch.autograd.Function):
    """The gradients behind ``_TransformBridge.backward``: the simulation run
    again under autograd on plain tensors (the adjoint kernels), in a
    Function's forward where no transform intercepts it."""

    @staticmethod
    def forward(simulator, keys, others, needs, grad, *tensors):
        with torch.enable_grad():
            leaves = [t.detach().requires_grad_(n and torch.is_floating_point(t)) for t, n in zip(tensors, needs)]
            out = simulator.simulate(**others, **dict(zip(keys, leaves)))
            wanted = [leaf for leaf in leaves if leaf.requires_grad]
            found = iter(torch.autograd.grad(out, wanted, grad, allow_unused=True) if wanted else ())
            grads = [next(found) if leaf.requires_grad else None for leaf in leaves]
        return tuple(torch.zeros_like(t) if n and g is None else g for t, n, g in zip(tensors, needs, grads))

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, *grads):
        raise NotImplementedError("a gradient is not differentiated again here")


def _directional(outputs: list, leaves: list, tangents: list) -> list:
    """J v for the map leaves -> outputs by a double backward (the maps are
    cheap torch operations; a None output or one the leaves do not reach
    gets None)."""
    live = [(i, out) for i, out in enumerate(outputs) if isinstance(out, torch.Tensor) and out.requires_grad]
    result: list = [None] * len(outputs)
    if not live or not leaves:
        return result
    probes = [torch.zeros_like(out, requires_grad=True) for _, out in live]
    grads = torch.autograd.grad([out for _, out in live], leaves, probes, create_graph=True, allow_unused=True)
    total = None
    for grad, tangent in zip(grads, tangents):
        if grad is None or not grad.requires_grad:
            continue
        term = (grad * tangent.to(grad.dtype)).sum()
        total = term if total is None else total + term
    if total is None:
        return result
    reach = torch.autograd.grad(total, probes, allow_unused=True)
    for (i, _), value in zip(live, reach):
        result[i] = value
    return result


class _BridgeTangent(torch.autograd.Function):
    """The tangent behind ``_TransformBridge.jvp``, a Function of its own so
    that its forward sees plain tensors."""

    @staticmethod
    def forward(simulator, keys, others, count, *both):
        from ..sequence import _complex

        primals, tangents = both[:count], both[count:]
        if simulator.signal_factor is None:
            raise NotImplementedError(
                f"torch.func.jvp of {type(simulator).__name__}, whose signal post-processes the samples: "
                "use torch.autograd.forward_ad (dual tensors) instead"
            )
        with torch.enable_grad():
            leaves, seeds = [], []
            values = dict(others)
            for name, primal, tangent in zip(keys, primals, tangents):
                if tangent is not None and torch.is_floating_point(primal):
                    primal = primal.detach().requires_grad_(True)
                    leaves.append(primal)
                    seeds.append(tangent)
                values[name] = primal
            held, sequence = simulator._split({**simulator.bound, **values})
            device, packed, _, tissue, states, _, _, _, dephasing = simulator._prepared(held, sequence)
            carried = [_native.TISSUE_INDEX[name] for name in _complex.GRADIENT_FIELDS]
            # the other properties too, to refuse a tangent that would be dropped
            others = [i for i, name in enumerate(_native.TISSUE_NAMES)
                      if name not in _complex.GRADIENT_FIELDS and tissue.buffers[i] is not None]
            targets = ([packed.duration, packed.flip, packed.phase] + [tissue.buffers[i] for i in carried]
                       + [tissue.buffers[i] for i in others])
            sample_factor = None if dephasing is None else dephasing.factor()
            if sample_factor is not None:
                targets += [sample_factor.real, sample_factor.imag]
            moved = _directional(targets, leaves, seeds)
        beside = moved[3 + len(carried):3 + len(carried) + len(others)]
        dropped = [_native.TISSUE_NAMES[i] for i, value in zip(others, beside)
                   if value is not None and bool((value != 0).any())]
        if dropped:
            raise NotImplementedError(f"forward mode carries no tangent along {dropped}")
        reached = moved[3:3 + len(carried)]
        for index, value in zip(carried, reached):
            if value is not None and tissue.buffers[index] is None:
                raise NotImplementedError("a tangent reached a tissue property the simulation leaves out")
        factor = simulator.signal_factor
        plain = dataclasses.replace(tissue, buffers=[None if b is None else b.detach() for b in tissue.buffers])
        tissue_tangents = {name: value for name, value in zip(_complex.GRADIENT_FIELDS, reached) if value is not None}
        profile = simulator._profile(sequence)
        if profile is not None or tissue.semisolid:
            from ..sequence import _semisolid

            # buffers laid out together all carry a tangent slot: keep the live ones
            live = {name: value for name, value in tissue_tangents.items() if bool((value != 0).any())}
            signal, tangent_states = _semisolid.tangent_of(packed, plain, profile, states=states,
                                                           tissue_tangents=live, factor=factor,
                                                           duration_tangent=moved[0], flip_tangent=moved[1],
                                                           phase_tangent=moved[2])
            both_out = tangent_states if sample_factor is None else (signal, tangent_states)
        else:
            both_out = _complex.jvp_complex(
                packed, plain, states=states, factor=factor, tissue_tangents=tissue_tangents,
                flip=packed.flip.detach(), phase=packed.phase.detach(), duration=packed.duration.detach(),
                duration_tangent=moved[0], flip_tangent=moved[1], phase_tangent=moved[2],
                with_signal=sample_factor is not None,
            )
        if sample_factor is None:
            tangent = both_out
        else:
            # d(f s) = f ds + s df, df along the same direction
            signal, tangent = both_out
            d_real, d_imag = moved[-2], moved[-1]
            d_factor = torch.complex(torch.zeros_like(sample_factor.real) if d_real is None else d_real,
                                     torch.zeros_like(sample_factor.imag) if d_imag is None else d_imag)
            tangent = sample_factor.detach() * tangent + d_factor * signal
        batch = simulator._batch(held)
        return tangent.reshape(*batch, tangent.shape[-1])

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, grad):
        raise NotImplementedError("a directional derivative is not differentiated again here")
