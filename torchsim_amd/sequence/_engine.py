"""Tissue preparation, the real-subspace verdict, and the native launches.

This is the host side of the reference's engine and dispatch
(sequence/_simulation.py:359-484 ``EpgEngine.simulate``, :661-737
``_prepare_tissue``; sequence/_accelerators.py:549-571 ``_one_axis``,
:574-746 ``simulate_native``, :1953-2017 the real-subspace verdict), reduced to
what the gfx950 library runs: one train per launch, single pool or one
exchanging pool, instantaneous pulses.
"""

from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch
from torch.autograd import forward_ad

from .. import _native
from ._events import INVERSION, RECORD, PackedEvents, resident

#: Relaxation times are clamped here before they are used as denominators
#: (reference sequence/_simulation.py:59, :734-736).
MINIMUM_RELAXATION_TIME_MS = 1e-6

#: How far from one axis a pulse or sample phase may sit and still count as on
#: it (reference sequence/_accelerators.py ``_PHASE_TOLERANCE``).
PHASE_TOLERANCE = 1e-5

#: The identity of each optional property: at it, the term is absent.
_IDENTITY = {
    "m0": 1.0,
    "b1": 1.0,
    "b1_phase_rad": 0.0,
    "b0_hz": 0.0,
    "inversion_efficiency": 1.0,
    "diffusion_um2_per_ms": 0.0,
    "velocity_m_per_s": 0.0,
    "bound_fraction": 0.0,
    "pool_b_fraction": 0.0,
}
_FEATURE_OF = {
    "m0": _native.FEATURE_DENSITY,
    "b1": _native.FEATURE_TRANSMIT,
    "inversion_efficiency": _native.FEATURE_INVERTING,
    "diffusion_um2_per_ms": _native.FEATURE_DIFFUSING,
    "b0_hz": _native.FEATURE_OFF_AXIS,
    "b1_phase_rad": _native.FEATURE_OFF_AXIS,
}


@dataclass
class PreparedTissue:
    """The 17 FP32 tissue buffers in ABI order, on the device."""

    buffers: list[torch.Tensor | None]
    shape: torch.Size
    atoms: int
    features: int
    stride: int  # 1: optional buffers hold one value per atom; 0: one in total
    exchanging: bool
    semisolid: bool = False  # a magnetization-transfer pool (bound fraction set)

    @property
    def device(self) -> torch.device:
        return self.buffers[0].device


def carries_derivative(value: Any) -> bool:
    """Whether a value carries a derivative: it requires grad, or it holds a
    forward-mode tangent (a dual number sits at its primal, which may be a
    property's identity, and its term must not be dropped there)."""
    if not isinstance(value, torch.Tensor):
        return False
    return value.requires_grad or forward_ad.unpack_dual(value).tangent is not None


def _at_identity(name: str, value: Any) -> bool:
    if name not in _IDENTITY:
        return False
    if isinstance(value, torch.Tensor):
        if carries_derivative(value):
            return False
        if value.numel() == 1:
            return float(value.reshape(())) == _IDENTITY[name]
        return False
    return float(value) == _IDENTITY[name]


def _on_device(value: Any, device: torch.device) -> torch.Tensor:
    """A property as a float32 tensor on ``device``; numbers and host scalars
    become a device fill (no host-to-device copy, so no wait on the stream)."""
    if isinstance(value, torch.Tensor):
        if value.device.type == "cpu" and value.numel() == 1 and not carries_derivative(value):
            return torch.full(value.shape, float(value.reshape(())), dtype=torch.float32, device=device)
        return value.to(device=device, dtype=torch.float32)
    if isinstance(value, (int, float, np.floating, np.integer)):
        return torch.full((), float(value), dtype=torch.float32, device=device)
    return torch.as_tensor(value, device=device).to(torch.float32)


def prepare_tissue(
    properties: dict[str, Any],
    device: torch.device,
    differentiated: tuple[str, ...] = (),
) -> PreparedTissue:
    """Lay the tissue out as the kernels read it.

    ``properties`` maps tissue field names (``t1_ms``, ``t2_ms``, ``m0``, ...)
    to numbers or tensors. A property left at its identity and not
    differentiated is absent: its feature bit stays off and its buffer is not
    read. Fields being differentiated are always laid out per atom.
    """
    given = {name: value for name, value in properties.items() if value is not None}
    for name in ("t1_ms", "t2_ms"):
        if name not in given:
            raise ValueError(f"{name} is required")
    features = 0
    live: dict[str, torch.Tensor] = {}
    for name, value in given.items():
        if name not in _native.TISSUE_INDEX:
            raise ValueError(f"unknown tissue property {name!r}")
        tensor = _on_device(value, device)
        if name in ("t1_ms", "t2_ms") or name in differentiated or not _at_identity(name, value):
            live[name] = tensor
            features |= _FEATURE_OF.get(name, 0)
    exchanging = "pool_b_fraction" in live
    semisolid = "bound_fraction" in live
    if semisolid:
        # the pool's exchange rate and T1 take the reference's defaults
        # (sequence/_simulation.py:212-214) where the caller gave none
        for name, default in (("bound_exchange_rate", 0.0), ("t1_bound_ms", 1000.0)):
            if name not in live:
                live[name] = _on_device(given.get(name, default), device)
    shape = torch.broadcast_shapes(*(value.shape for value in live.values()))
    atoms = 1
    for size in shape:
        atoms *= int(size)
    optional = [name for name in live if name not in ("t1_ms", "t2_ms")]
    narrowed = all(live[name].numel() == 1 and name not in differentiated for name in optional)
    stride = 0 if narrowed else 1
    buffers: list[torch.Tensor | None] = [None] * _native.TISSUE_COUNT
    for name, tensor in live.items():
        if stride == 0 and name in optional:
            flat = tensor.reshape(-1)[:1].contiguous()
        else:
            flat = tensor.expand(shape).reshape(-1).contiguous()
        if name in ("t1_ms", "t2_ms", "t1_bound_ms", "t1_pool_b_ms", "t2_pool_b_ms"):
            flat = flat.clamp_min(MINIMUM_RELAXATION_TIME_MS)
        buffers[_native.TISSUE_INDEX[name]] = flat
    # Buffers a live feature reads but the caller left unset take their identity.
    for name, bit in _FEATURE_OF.items():
        index = _native.TISSUE_INDEX[name]
        if features & bit and buffers[index] is None:
            fill = torch.full((1 if stride == 0 else atoms,), _IDENTITY[name], dtype=torch.float32, device=device)
            buffers[index] = fill
    return PreparedTissue(buffers, torch.Size(shape), atoms, features, stride, exchanging, semisolid)


def half_turns(packed: PackedEvents) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor | None] | None:
    """Bring pulses and samples onto one phase axis, or say it cannot be done.

    Returns ``(phase_offsets, flip_signs, sample_signs)`` where subtracting a
    whole number of half turns from a phase is compensated by a sign on the
    flip (pulses) or on the recorded sample (samples); ``None`` if some pulse
    or sample sits off that axis (reference sequence/_accelerators.py:500-546).
    """
    # a few thousand events: NumPy on the host, no torch dispatch per op
    kind = packed.kind.cpu().numpy()
    action = packed.action.cpu().numpy().astype(np.int32)
    phase = packed.phase.detach().cpu().numpy().astype(np.float64)
    turning = (kind == 1) & ((action & INVERSION) == 0)
    recording = (kind == 2) & ((action & RECORD) != 0)
    if not turning.any():
        return None
    anchor = recording if recording.any() else turning
    delta = phase - phase[anchor][0]
    on_axis = turning | recording
    out_by = np.where(on_axis, np.round(delta / np.pi), 0.0)
    residue = np.where(on_axis, delta - out_by * np.pi, 0.0)
    if float(np.abs(residue).max()) > PHASE_TOLERANCE:
        return None
    # cos(k pi) for a whole number of half turns k: +1 for even k, -1 for odd
    sign = np.where(np.fmod(out_by, 2.0) == 0.0, 1.0, -1.0)
    flips = np.where(turning, sign, 1.0)
    samples = torch.from_numpy(sign[recording]) if recording.any() else None
    return torch.from_numpy(out_by), torch.from_numpy(flips), samples


def real_subspace(packed: PackedEvents, tissue: PreparedTissue,
                  device: torch.device | None = None) -> tuple[bool, torch.Tensor | None, torch.Tensor | None]:
    """Whether the real-subspace kernel applies; the signed flips and the
    sample signs if so (None where no sign flips). The verdict is worked out
    once per packed stream; with ``device``, the flips and signs come back
    resident there (uploaded once, asynchronously)."""
    if tissue.exchanging or tissue.semisolid or tissue.features & (_native.FEATURE_OFF_AXIS | _native.FEATURE_MOVING):
        return False, None, None
    verdict = packed.__dict__.get("_real_subspace")
    if verdict is None:
        turns = half_turns(packed)
        if turns is None:
            verdict = (False, None, None)
        else:
            _, flips, samples = turns
            flip = packed.flip.detach().cpu() * flips.to(torch.float32) if bool((flips != 1).any()) else None
            packed.__dict__["_flip_signs"] = flips.to(torch.float32) if flip is not None else None
            if samples is not None and bool((samples == 1).all()):
                samples = None
            if samples is not None:
                samples = samples.to(torch.float32)
            verdict = (True, flip, samples)
        packed.__dict__["_real_subspace"] = verdict
    real, flip, samples = verdict
    if real and device is not None and (flip is not None or samples is not None):
        flip, samples = resident(packed, "signs", [flip, samples], device)
    return real, flip, samples


_COLUMN_INDEX = {
    "t1_ms": 0,
    "t2_ms": 1,
    "m0": 2,
    "b1": 3,
    "inversion_efficiency": 6,
}


def simulate_real(
    packed: PackedEvents,
    tissue: PreparedTissue,
    *,
    states: int,
    columns: tuple[str, ...] = (),
    factor: complex = 1.0,
    flip: torch.Tensor | None = None,
    samples: torch.Tensor | None = None,
    tile: int | str = "auto",
    duration: torch.Tensor | None = None,
) -> tuple[torch.Tensor, torch.Tensor | None]:
    """Signal and (optionally) Jacobian columns in one real-subspace launch.

    Returns complex64 ``(atoms, outputs)`` and ``(atoms, len(columns), outputs)``
    already multiplied by ``factor`` (and by the sample signs). ``tile`` is a
    state-tile index, -1 for the library's cost model, or ``"auto"`` for the
    measured choice (``_tuning.tile_for``, probed once per configuration).
    """
    if tile == "auto":
        # measuring synchronizes, which a stream under graph capture cannot do
        if torch.cuda.is_current_stream_capturing():
            tile = -1
        else:
            tile = _tuned_tile(packed, tissue, states=states, columns=columns, factor=factor, flip=flip,
                               samples=samples, duration=duration)
    device = tissue.device
    _native.require_device(device)
    lib = _native.load()
    events = list(packed.on(device))
    if flip is not None:
        events[2] = flip.detach().to(device=device, dtype=torch.float32).contiguous()
    if duration is not None:
        events[0] = duration.detach().to(device=device, dtype=torch.float32).contiguous()
    atoms, outputs = tissue.atoms, packed.output_count
    signal = torch.empty((atoms, outputs), dtype=torch.complex64, device=device)
    jacobian = None
    if columns:
        jacobian = torch.empty((atoms, len(columns), outputs), dtype=torch.complex64, device=device)
    if atoms == 0 or outputs == 0:
        return signal, jacobian
    ptrs = _native.pointers([*tissue.buffers, *events])
    cols = (ctypes.c_int32 * max(1, len(columns)))(*[_COLUMN_INDEX[name] for name in columns])
    sign_buffer = None
    if samples is not None:
        sign_buffer = samples.to(device=device, dtype=torch.float32).contiguous()
    with torch.cuda.device(device):
        status = lib.tsim_simulate_jacobian(
            ptrs,
            atoms,
            packed.event_count,
            states,
            outputs,
            tissue.features,
            tissue.stride,
            cols,
            len(columns),
            float(complex(factor).real),
            float(complex(factor).imag),
            None if sign_buffer is None else sign_buffer.data_ptr(),
            signal.data_ptr(),
            None if jacobian is None else jacobian.data_ptr(),
            tile,
            _native.stream_of(device),
        )
    _native.check(status)
    return signal, jacobian


#: Basis directions one projected launch carries (4 per lane, 4 lanes per atom).
PROJECTED_RANK_MAX = 16


def projects_fused(tissue: PreparedTissue, columns: tuple[str, ...], rank: int) -> bool:
    """Whether ``simulate_projected`` takes this launch: no diffusion, no
    inversion-efficiency column, at most PROJECTED_RANK_MAX directions."""
    return (rank <= PROJECTED_RANK_MAX and not tissue.features & _native.FEATURE_DIFFUSING
            and "inversion_efficiency" not in columns)


def simulate_projected(
    packed: PackedEvents,
    tissue: PreparedTissue,
    basis: torch.Tensor,
    *,
    states: int,
    columns: tuple[str, ...] = (),
    factor: complex = 1.0,
    flip: torch.Tensor | None = None,
    samples: torch.Tensor | None = None,
    tile: int = -1,
) -> tuple[torch.Tensor, torch.Tensor | None]:
    """``simulate_real`` with the samples contracted with ``basis``
    ``(outputs, rank)`` in the kernel: complex64 ``(atoms, rank)`` and
    ``(atoms, len(columns), rank)``, i.e. ``signal @ basis`` and
    ``jacobian @ basis`` without either reaching memory."""
    device = tissue.device
    _native.require_device(device)
    lib = _native.load()
    outputs = packed.output_count
    if basis.ndim != 2 or basis.shape[0] != outputs:
        raise ValueError(f"the basis must be (samples={outputs}, rank), got {tuple(basis.shape)}")
    rank = int(basis.shape[1])
    atoms = tissue.atoms
    if outputs == 0:
        signal = torch.zeros((atoms, rank), dtype=torch.complex64, device=device)
        jacobian = torch.zeros((atoms, len(columns), rank), dtype=torch.complex64, device=device) if columns else None
        return signal, jacobian
    events = list(packed.on(device))
    if flip is not None:
        events[2] = flip.detach().to(device=device, dtype=torch.float32).contiguous()
    held = basis.detach().to(device=device, dtype=torch.complex64).contiguous()
    signal = torch.empty((atoms, rank), dtype=torch.complex64, device=device)
    jacobian = torch.empty((atoms, len(columns), rank), dtype=torch.complex64, device=device) if columns else None
    if atoms == 0:
        return signal, jacobian
    ptrs = _native.pointers([*tissue.buffers, *events])
    cols = (ctypes.c_int32 * max(1, len(columns)))(*[_COLUMN_INDEX[name] for name in columns])
    sign_buffer = None if samples is None else samples.to(device=device, dtype=torch.float32).contiguous()
    with torch.cuda.device(device):
        status = lib.tsim_simulate_projected(
            ptrs, atoms, packed.event_count, states, outputs, tissue.features, tissue.stride, cols, len(columns),
            float(complex(factor).real), float(complex(factor).imag),
            None if sign_buffer is None else sign_buffer.data_ptr(), held.data_ptr(), rank, signal.data_ptr(),
            None if jacobian is None else jacobian.data_ptr(), tile, _native.stream_of(device),
        )
    _native.check(status)
    return signal, jacobian


def narrow_tissue(tissue: PreparedTissue, atoms: int) -> PreparedTissue:
    """The first ``atoms`` atoms of a prepared tissue (views, no copy)."""
    atoms = min(atoms, tissue.atoms)
    buffers = []
    for index, buffer in enumerate(tissue.buffers):
        per_atom = buffer is not None and (index in (0, 1) or tissue.stride == 1)
        buffers.append(buffer[:atoms] if per_atom else buffer)
    return dataclasses.replace(tissue, buffers=buffers, shape=torch.Size([atoms]), atoms=atoms)


def _tuned_tile(packed: PackedEvents, tissue: PreparedTissue, *, states: int, columns: tuple[str, ...],
                factor: complex, flip: torch.Tensor | None, samples: torch.Tensor | None,
                duration: torch.Tensor | None) -> int:
    from . import _tuning

    probe_tissue = narrow_tissue(tissue, _tuning.PROBE_ATOMS)
    if probe_tissue.atoms == 0 or packed.output_count == 0:
        return -1

    def probe(candidate: int) -> None:
        simulate_real(packed, probe_tissue, states=states, columns=columns, factor=factor, flip=flip,
                      samples=samples, tile=candidate, duration=duration)

    return _tuning.tile_for(tissue.device, states, tuple(columns), tissue.features, probe)


#: Tissue fields the real-subspace adjoint returns gradients for (the
#: reference's simulate_vjp writes all 17; off-resonance, transmit phase,
#: diffusion and the pools take the complex path, which has no adjoint here).
GRADIENT_FIELDS = ("t1_ms", "t2_ms", "m0", "b1", "inversion_efficiency")


def _graph_leaves(tensor: torch.Tensor) -> list[torch.Tensor]:
    """The leaf tensors requiring grad that ``tensor``'s graph reaches."""
    # the visited nodes are held so that their ids stay unique
    leaves, held, seen, todo = [], [], set(), [tensor.grad_fn]
    while todo:
        node = todo.pop()
        if node is None or id(node) in seen:
            continue
        held.append(node)
        seen.add(id(node))
        variable = getattr(node, "variable", None)
        if variable is not None:
            leaves.append(variable)
        todo.extend(parent for parent, _ in node.next_functions)
    return leaves


def zero_intervals_reached(packed: PackedEvents) -> np.ndarray:
    """Which events have a zero interval whose value the caller's graph
    reaches (a TI of 0 that requires grad, say): those need a duration
    gradient although their interval is empty. A packer that knows says so
    on the stream; otherwise they are the nonzero entries of J w, J the
    Jacobian of the durations with respect to the graph's leaves and w
    positive random weights, found by one double backward through the
    packing; if that cannot run, every zero interval is taken."""
    known = packed.__dict__.get("_zero_intervals")
    if known is not None:
        return known
    duration = packed.duration
    values = duration.detach()
    zero = (values == 0).reshape(-1).cpu().numpy()
    if not duration.requires_grad or not zero.any():
        return np.zeros(values.numel(), dtype=bool)
    leaves = _graph_leaves(duration)
    if not leaves:
        return np.zeros(values.numel(), dtype=bool)
    probe = torch.zeros_like(duration, requires_grad=True)
    try:
        grads = torch.autograd.grad(duration, leaves, grad_outputs=probe, create_graph=True, allow_unused=True)
        generator = torch.Generator().manual_seed(0)
        total = None
        for grad in grads:
            if grad is None or not grad.requires_grad:
                continue
            weight = 0.5 + torch.rand(grad.shape, generator=generator, dtype=torch.float64)
            term = (grad * weight.to(device=grad.device, dtype=grad.dtype)).sum()
            total = term if total is None else total + term
        if total is None:
            return np.zeros(values.numel(), dtype=bool)
        (reach,) = torch.autograd.grad(total, probe, retain_graph=True, allow_unused=True)
    except RuntimeError:
        return zero
    if reach is None:
        return np.zeros(values.numel(), dtype=bool)
    return zero & (reach.detach().reshape(-1).cpu().numpy() != 0)


def op_count(packed: PackedEvents, zero_intervals: np.ndarray | None = None) -> int:
    """How many macro-ops the library groups the stream into: each event's
    parts -- its interval (position 0, if nonzero), a pre-shift (3), its
    rotation or inversion (1) or record (2), a post-shift (3), a spoil (1) or
    a shift after it (3) -- in order, a new op starting wherever a part's
    position does not exceed its predecessor's or follows a shift (the rule
    the op compiler applies on the device)."""
    kind = packed.kind.cpu().numpy()
    action = packed.action.cpu().numpy().astype(np.int32)
    timed = packed.duration.detach().cpu().numpy() != 0
    if zero_intervals is not None:
        timed = timed | zero_intervals
    position = np.full((kind.size, 5), -1, dtype=np.int8)
    position[timed, 0] = 0
    position[(action & 1) != 0, 1] = 3
    position[kind == 1, 2] = 1
    position[(kind == 2) & ((action & RECORD) != 0), 2] = 2
    position[(action & 2) != 0, 3] = 3
    position[(action & 16) != 0, 4] = 3
    position[(action & 8) != 0, 4] = 1
    parts = position.reshape(-1)
    parts = parts[parts >= 0]
    if parts.size == 0:
        return 0
    return 1 + int(np.count_nonzero((parts[1:] <= parts[:-1]) | (parts[:-1] == 3)))


_NO_ZERO_INTERVALS: dict = {}


def _no_zero_intervals(count: int, device: torch.device) -> torch.Tensor:
    """A resident all-zero duration mask of at least ``count`` bytes."""
    key = str(device)
    held = _NO_ZERO_INTERVALS.get(key)
    if held is None or held.numel() < count:
        held = torch.zeros(max(count, 4096), dtype=torch.uint8, device=device)
        _NO_ZERO_INTERVALS[key] = held
    return held


def adjoint_real(
    packed: PackedEvents,
    tissue: PreparedTissue,
    *,
    states: int,
    columns: tuple[str, ...],
    flip: torch.Tensor,
    duration: torch.Tensor,
    signal_cotangent: torch.Tensor | None,
    jacobian_cotangent: torch.Tensor | None,
    checkpoint_pulses: int = 64,
    tissue_wanted: tuple[str, ...] = (),
    op_bound: int | None = None,
    zero_intervals: np.ndarray | bool | None = None,
) -> tuple[torch.Tensor, torch.Tensor, dict[str, torch.Tensor]]:
    """d(loss)/d(flip) and d(loss)/d(duration), one entry per event, and
    d(loss)/d(field) per atom for each field of ``tissue_wanted`` (among
    GRADIENT_FIELDS), from the loss's real cotangents on the recorded m0*a0
    (atoms, outputs) and on the Jacobian columns (atoms, P, outputs). The
    state is checkpointed every ``checkpoint_pulses`` macro-ops (pulses of an
    MRF train) and recomputed per segment. ``op_bound`` sizes the workspace
    (default: the stream's op count, worked out on the host; -1 has the
    library read it back from the device). ``zero_intervals`` flags the
    events with a zero interval whose duration gradient is wanted (None:
    none of them, True: all of them, the reference's convention; every timed
    event gets its own)."""
    device = tissue.device
    _native.require_device(device)
    if states < 2:
        # a stream without shifts never fills order 1, so two orders give the
        # one-order answer exactly
        shifts = (packed.action.cpu().numpy().astype(np.int32) & (1 | 2 | 16)) != 0
        if shifts.any():
            raise NotImplementedError("the real-subspace adjoint carries at least 2 orders and this stream shifts")
        states = 2
    lib = _native.load()
    events = list(packed.on(device))
    events[2] = flip.detach().to(device=device, dtype=torch.float32).contiguous()
    events[0] = duration.detach().to(device=device, dtype=torch.float32).contiguous()
    count = packed.event_count
    grad_flip = torch.empty(count, dtype=torch.float32, device=device)
    grad_duration = torch.empty(count, dtype=torch.float32, device=device)
    grads = {}
    slots = (ctypes.c_void_p * _native.TISSUE_COUNT)()
    for name in tissue_wanted:
        if name not in GRADIENT_FIELDS:
            raise NotImplementedError(f"the real-subspace adjoint returns no gradient for {name!r}")
        grads[name] = torch.empty(tissue.atoms, dtype=torch.float32, device=device)
        slots[_native.TISSUE_INDEX[name]] = grads[name].data_ptr()
    ptrs = _native.pointers([*(None if b is None else b.detach() for b in tissue.buffers), *events])
    cols = (ctypes.c_int32 * max(1, len(columns)))(*[_COLUMN_INDEX[name] for name in columns])
    ws = None if signal_cotangent is None else signal_cotangent.to(device=device, dtype=torch.float32).contiguous()
    wj = None if jacobian_cotangent is None else jacobian_cotangent.to(device=device, dtype=torch.float32).contiguous()
    every = zero_intervals is True
    if every:
        zero_intervals = packed.duration.detach().cpu().numpy() == 0
    if zero_intervals is None or not zero_intervals.any():
        mask = _no_zero_intervals(count, device)
    else:
        mask = torch.from_numpy(zero_intervals.astype(np.uint8)).to(device)
    with torch.cuda.device(device):
        status = lib.tsim_simulate_vjp_jacobian(
            ptrs,
            tissue.atoms,
            count,
            states,
            packed.output_count,
            tissue.features,
            tissue.stride,
            cols,
            len(columns),
            None if ws is None else ws.data_ptr(),
            None if wj is None else wj.data_ptr(),
            checkpoint_pulses,
            op_count(packed, zero_intervals) if op_bound is None else op_bound,
            grad_flip.data_ptr(),
            grad_duration.data_ptr(),
            slots,
            _native.TISSUE_COUNT if grads else 0,
            None if every else mask.data_ptr(),
            _native.stream_of(device),
        )
    _native.check(status)
    return grad_flip, grad_duration, grads


class FusedRealJacobian(torch.autograd.Function):
    """Signal and Jacobian of the real subspace, differentiable in the event
    flips and durations and in the per-atom T1, T2, M0, B1 and inversion
    efficiency (reference sequence/_accelerators.py:1402-1858: the forward of
    ``_NativeEpg``/``_NativeEpgJvp`` and their backward, here one fused
    forward kernel and one checkpointed adjoint kernel). The tissue inputs
    are the prepared per-atom buffers the kernels read, so autograd carries
    their gradients back through the layout to the caller's maps."""

    @staticmethod
    def forward(ctx, flip, duration, t1, t2, m0, b1, ie, packed, tissue, states, columns, factor, samples,
                checkpoint_pulses, zero_intervals):
        signal, jacobian = simulate_real(
            packed, tissue, states=states, columns=columns, factor=factor, flip=flip,
            samples=samples, duration=duration,
        )
        ctx.save_for_backward(flip, duration)
        ctx.shapes = [None if v is None else v.shape for v in (t1, t2, m0, b1, ie)]
        ctx.packed = packed
        ctx.tissue = tissue
        ctx.states = states
        ctx.columns = columns
        ctx.factor = complex(factor)
        ctx.samples = samples
        ctx.checkpoint_pulses = checkpoint_pulses
        ctx.zero_intervals = zero_intervals
        if jacobian is None:
            jacobian = torch.empty((tissue.atoms, 0, packed.output_count), dtype=torch.complex64, device=signal.device)
        return signal, jacobian

    @staticmethod
    def backward(ctx, grad_signal, grad_jacobian):
        flip, duration = ctx.saved_tensors
        nothing = (None,) * 15
        # every output is c * v with v real and c = factor * i (* sample sign):
        # the loss's derivative along v is Re(conj(c) * grad).
        c = torch.tensor(ctx.factor * 1j, dtype=torch.complex64, device=flip.device)
        scale = c.conj()
        if ctx.samples is not None:
            scale = scale * ctx.samples.to(device=flip.device, dtype=torch.complex64)
        ws = None if grad_signal is None else (grad_signal * scale).real
        wj = None
        if grad_jacobian is not None and len(ctx.columns):
            wj = (grad_jacobian * scale).real
        if ws is None and wj is None:
            return nothing
        wanted = tuple(name for name, need in zip(GRADIENT_FIELDS, ctx.needs_input_grad[2:7]) if need)
        grad_flip, grad_duration, grads = adjoint_real(
            ctx.packed, ctx.tissue, states=ctx.states, columns=ctx.columns, flip=flip, duration=duration,
            signal_cotangent=ws, jacobian_cotangent=wj, checkpoint_pulses=ctx.checkpoint_pulses,
            tissue_wanted=wanted, zero_intervals=ctx.zero_intervals if ctx.needs_input_grad[1] else None,
        )
        tissue_out = []
        for name, shape in zip(GRADIENT_FIELDS, ctx.shapes):
            g = grads.get(name)
            if g is not None and shape is not None and g.numel() != torch.Size(shape).numel():
                g = g.sum().reshape(shape)  # one value for every atom
            tissue_out.append(g)
        return (
            grad_flip.to(flip.dtype) if ctx.needs_input_grad[0] else None,
            grad_duration.to(duration.dtype) if ctx.needs_input_grad[1] else None,
            *tissue_out,
            None, None, None, None, None, None, None, None,
        )
