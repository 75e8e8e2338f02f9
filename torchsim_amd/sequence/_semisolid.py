"""Launch of the semisolid-pool (magnetization-transfer) and slice-profiled
pulse kernel.

Reference: the SEMISOLID pool kind and the PROFILED rotation of the complex
state machine (src/torchsim/_epg_cpu.cpp:4216-4566 with pool kind 1,
2041-2049, 2107-2160, 2273-2304) as ``simulate_native`` reaches them: an
absorption table sized by ``_absorption_table`` (sequence/_simulation.py:
80-94), every voxel copied once per slice position and its signal averaged
over them (sequence/_accelerators.py:144-158, 730-734).
"""

from __future__ import annotations

import ctypes
import dataclasses
import threading

import torch

from .. import _native
from . import _complex
from ._engine import PreparedTissue
from ._events import PackedEvents
from ._lineshape import Lineshape, lineshape_reaching
from ._transition import TransitionTable

_COLUMN_INDEX = {"t1_ms": 0, "t2_ms": 1, "m0": 2, "b1": 3}
#: The tissue maps derivatives reach through these kernels (their columns).
CARRIED = tuple(_COLUMN_INDEX)
_TABLES: dict = {}
_TABLES_LOCK = threading.Lock()


def _resident(table: Lineshape, device: torch.device) -> torch.Tensor:
    """The table's (bins, 2) float32 knots on ``device``, uploaded once."""
    key = (table.bins, table.step, str(device))
    with _TABLES_LOCK:
        held = _TABLES.get(key)
        if held is None:
            held = _TABLES[key] = table.packed(device)
    return held


def absorption(packed: PackedEvents, tissue: PreparedTissue) -> Lineshape:
    """The lineshape reaching the largest |pulse frequency - off-resonance| of the run."""
    pulses = float(packed.rf_frequency_hz.detach().abs().max()) if packed.event_count else 0.0
    b0 = tissue.buffers[_native.TISSUE_INDEX["b0_hz"]]
    voxel = float(b0.detach().abs().max()) if b0 is not None and b0.numel() else 0.0
    return lineshape_reaching(pulses + voxel)


def simulate_semisolid(
    packed: PackedEvents,
    tissue: PreparedTissue,
    *,
    states: int,
    columns: tuple[str, ...] = (),
    factor: complex = 1.0,
    flip: torch.Tensor | None = None,
    phase: torch.Tensor | None = None,
    duration: torch.Tensor | None = None,
) -> tuple[torch.Tensor, torch.Tensor | None]:
    """Complex64 signal (atoms, outputs) and Jacobian (atoms, P, outputs) of a
    tissue with a semisolid pool; columns among t1_ms, t2_ms, m0, b1; the
    events' flips, phases and durations from the given tensors where given."""
    device = tissue.device
    _native.require_device(device)
    for name in columns:
        if name not in _COLUMN_INDEX:
            raise NotImplementedError(f"the semisolid kernel carries no tangent along {name!r}")
    if tissue.exchanging:
        raise NotImplementedError("a semisolid pool beside an exchanging pool (three pools) is not built")
    lib = _native.load()
    table = absorption(packed, tissue)
    knots = _resident(table, device)
    atoms, outputs = tissue.atoms, packed.output_count
    signal = torch.zeros((atoms, outputs), dtype=torch.complex64, device=device)
    jacobian = torch.zeros((atoms, len(columns), outputs), dtype=torch.complex64, device=device) if columns else None
    if atoms == 0 or outputs == 0:
        return signal, jacobian
    events = _complex._events(packed, device, flip=flip, phase=phase, duration=duration)
    ptrs = _native.pointers([*(None if b is None else b.detach() for b in tissue.buffers), *events])
    cols = (ctypes.c_int32 * max(1, len(columns)))(*[_COLUMN_INDEX[name] for name in columns])
    with torch.cuda.device(device):
        status = lib.tsim_semisolid_jacobian(
            ptrs, atoms, packed.event_count, states, outputs, tissue.features, tissue.stride, cols, len(columns),
            float(complex(factor).real), float(complex(factor).imag), knots.data_ptr(), table.bins,
            float(table.step), signal.data_ptr(), None if jacobian is None else jacobian.data_ptr(),
            _native.stream_of(device),
        )
    _native.check(status)
    return signal, jacobian


def _across(tissue: PreparedTissue, locations: int) -> PreparedTissue:
    """Each atom copied once per slice position, voxel-major (buffers holding
    one value for every atom stay single)."""
    if locations == 1:
        return tissue
    buffers = [None if b is None else (b.repeat_interleave(locations) if b.numel() == tissue.atoms else b)
               for b in tissue.buffers]
    return dataclasses.replace(tissue, buffers=buffers, atoms=tissue.atoms * locations,
                               shape=torch.Size((tissue.atoms * locations,)))


def simulate_profiled(
    packed: PackedEvents,
    tissue: PreparedTissue,
    table: TransitionTable,
    *,
    states: int,
    columns: tuple[str, ...] = (),
    factor: complex = 1.0,
    flip: torch.Tensor | None = None,
    phase: torch.Tensor | None = None,
    duration: torch.Tensor | None = None,
) -> tuple[torch.Tensor, torch.Tensor | None]:
    """Signal (atoms, outputs) and Jacobian (atoms, P, outputs), complex64,
    with every pulse but an inversion played as the shaped pulse tabulated in
    ``table``: each atom simulated at every tabulated slice position and
    averaged (a semisolid pool runs too where the tissue has one)."""
    device = tissue.device
    _native.require_device(device)
    for name in columns:
        if name not in _COLUMN_INDEX:
            raise NotImplementedError(f"the profiled kernel carries no tangent along {name!r}")
    if tissue.exchanging:
        raise NotImplementedError("an exchanging pool is not built beside a profiled pulse")
    lib = _native.load()
    locations = table.points
    spread = _across(tissue, locations)
    knots = table.packed(device)
    lineshape = absorption(packed, tissue) if tissue.semisolid else None
    shape_knots = _resident(lineshape, device) if lineshape is not None else None
    atoms, outputs = spread.atoms, packed.output_count
    signal = torch.zeros((atoms, outputs), dtype=torch.complex64, device=device)
    jacobian = torch.zeros((atoms, len(columns), outputs), dtype=torch.complex64, device=device) if columns else None
    if atoms == 0 or outputs == 0:
        voxels = tissue.atoms
        return (signal.reshape(voxels, locations, outputs).mean(1),
                None if jacobian is None else jacobian.reshape(voxels, locations, len(columns), outputs).mean(1))
    events = _complex._events(packed, device, flip=flip, phase=phase, duration=duration)
    ptrs = _native.pointers([*(None if b is None else b.detach() for b in spread.buffers), *events])
    cols = (ctypes.c_int32 * max(1, len(columns)))(*[_COLUMN_INDEX[name] for name in columns])
    with torch.cuda.device(device):
        status = lib.tsim_profiled_jacobian(
            ptrs, atoms, packed.event_count, states, outputs, spread.features, spread.stride, cols, len(columns),
            float(complex(factor).real), float(complex(factor).imag), knots.data_ptr(), locations, table.bins,
            float(table.step), None if shape_knots is None else shape_knots.data_ptr(),
            0 if lineshape is None else lineshape.bins, 1.0 if lineshape is None else float(lineshape.step),
            signal.data_ptr(), None if jacobian is None else jacobian.data_ptr(), _native.stream_of(device),
        )
    _native.check(status)
    voxels = tissue.atoms
    signal = signal.reshape(voxels, locations, outputs).mean(1)
    if jacobian is not None:
        jacobian = jacobian.reshape(voxels, locations, len(columns), outputs).mean(1)
    return signal, jacobian


def simulate_shaped(
    packed: PackedEvents,
    tissue: PreparedTissue,
    table: TransitionTable | None,
    *,
    states: int,
    columns: tuple[str, ...] = (),
    factor: complex = 1.0,
    flip: torch.Tensor | None = None,
    phase: torch.Tensor | None = None,
    duration: torch.Tensor | None = None,
) -> tuple[torch.Tensor, torch.Tensor | None]:
    """:func:`simulate_profiled` where ``table`` names a shaped pulse, else
    :func:`simulate_semisolid`."""
    events = {"flip": flip, "phase": phase, "duration": duration}
    if table is not None:
        return simulate_profiled(packed, tissue, table, states=states, columns=columns, factor=factor, **events)
    return simulate_semisolid(packed, tissue, states=states, columns=columns, factor=factor, **events)


def refuse_uncarried(names) -> None:
    """Derivatives through these kernels travel along the tissue maps they
    write columns for.

    Raises
    ------
    NotImplementedError
        If ``names`` holds anything else.
    """
    others = [name for name in names if name not in CARRIED]
    if others:
        raise NotImplementedError(
            f"through a semisolid pool or a profiled pulse derivatives reach {', '.join(CARRIED)}; "
            f"{others} also carry one"
        )


class ColumnSignal(torch.autograd.Function):
    """The signal of a semisolid-pool or profiled-pulse run as a function of
    the tissue maps it was differentiated along, through the Jacobian columns
    the kernel wrote beside it: ``backward`` contracts the cotangent with the
    columns (grad_k = Re sum_n g_n conj(J_kn)) and ``jvp`` sums the columns
    times the tangents (the reference's ``_NativeEpg.backward`` / ``.jvp``,
    sequence/_accelerators.py:1511-1586, for these pools' tissue inputs). The
    signal and columns come in computed; the buffers are the maps' prepared
    per-atom values, in ``names`` order."""

    @staticmethod
    def forward(signal, jacobian, names, *buffers):
        return signal.clone()

    @staticmethod
    def setup_context(ctx, inputs, output):
        _, jacobian, names, *buffers = inputs
        ctx.save_for_backward(jacobian)
        ctx.save_for_forward(jacobian)
        ctx.shapes = [None if b is None else b.shape for b in buffers]

    @staticmethod
    def backward(ctx, grad):
        (jacobian,) = ctx.saved_tensors
        grads = []
        for k, (shape, need) in enumerate(zip(ctx.shapes, ctx.needs_input_grad[3:])):
            if not need or shape is None:
                grads.append(None)
                continue
            g = (grad * jacobian[:, k].conj()).real.sum(-1)
            # a buffer holding one value for every atom gathers every atom's term
            grads.append(g.reshape(shape) if g.numel() == shape.numel() else g.sum().reshape(shape))
        return (None, None, None, *grads)

    @staticmethod
    def jvp(ctx, _d_signal, _d_jacobian, _names, *tangents):
        (jacobian,) = ctx.saved_tensors
        total = None
        for k, tangent in enumerate(tangents):
            if tangent is None:
                continue
            term = jacobian[:, k] * tangent.reshape(-1, 1).to(jacobian.real.dtype)
            total = term if total is None else total + term
        return torch.zeros_like(jacobian[:, 0]) if total is None else total


def carried_signal(
    packed: PackedEvents,
    tissue: PreparedTissue,
    table: TransitionTable | None,
    *,
    states: int,
    carried: tuple[str, ...],
    factor: complex = 1.0,
) -> torch.Tensor:
    """The signal with reverse- and forward-mode derivatives along the tissue
    maps ``carried`` names (among :data:`CARRIED`), through their columns."""
    refuse_uncarried(carried)
    plain = dataclasses.replace(tissue, buffers=[None if b is None else b.detach() for b in tissue.buffers])
    signal, jacobian = simulate_shaped(packed, plain, table, states=states, columns=carried, factor=factor)
    buffers = [tissue.buffers[_native.TISSUE_INDEX[name]] for name in carried]
    return ColumnSignal.apply(signal, jacobian, carried, *buffers)


def event_tangent(
    packed: PackedEvents,
    tissue: PreparedTissue,
    table: TransitionTable | None,
    *,
    states: int,
    factor: complex,
    flip_tangent: torch.Tensor | None = None,
    phase_tangent: torch.Tensor | None = None,
    duration_tangent: torch.Tensor | None = None,
    flip: torch.Tensor | None = None,
    phase: torch.Tensor | None = None,
    duration: torch.Tensor | None = None,
) -> torch.Tensor:
    """The signal's directional derivative (voxels, outputs) along per-event
    tangents of the flips, phases and durations, through the semisolid pool
    and/or slice-profiled pulses (``table``), by ``tsim_semisolid_event_jvp``."""
    device = tissue.device
    _native.require_device(device)
    if tissue.exchanging:
        raise NotImplementedError("a semisolid pool beside an exchanging pool (three pools) is not built")
    lib = _native.load()
    lineshape = absorption(packed, tissue) if tissue.semisolid else None
    knots = _resident(lineshape, device) if lineshape is not None else None
    events = _complex._events(packed, device, flip=flip, phase=phase, duration=duration)

    def on_device(value):
        return None if value is None else value.detach().to(device=device, dtype=torch.float32).contiguous()

    directions = [on_device(v) for v in (flip_tangent, phase_tangent, duration_tangent)]
    locations = 1 if table is None else table.points
    spread = tissue if table is None else _across(tissue, locations)
    out = torch.zeros((spread.atoms, packed.output_count), dtype=torch.complex64, device=device)
    profile = None if table is None else table.packed(device)
    ptrs = _native.pointers([*(None if b is None else b.detach() for b in spread.buffers), *events])
    with torch.cuda.device(device):
        status = lib.tsim_semisolid_event_jvp(
            ptrs, spread.atoms, packed.event_count, states, packed.output_count, spread.features, spread.stride,
            float(complex(factor).real), float(complex(factor).imag),
            None if profile is None else profile.data_ptr(), locations, 0 if table is None else table.bins,
            1.0 if table is None else float(table.step), None if knots is None else knots.data_ptr(),
            0 if lineshape is None else lineshape.bins, 1.0 if lineshape is None else float(lineshape.step),
            *(None if d is None else d.data_ptr() for d in directions), out.data_ptr(), _native.stream_of(device))
    _native.check(status)
    if locations > 1:
        out = out.reshape(tissue.atoms, locations, -1).mean(1)
    return out


def tangent_of(
    packed: PackedEvents,
    tissue: PreparedTissue,
    table: TransitionTable | None,
    *,
    states: int,
    tissue_tangents: dict,
    factor: complex = 1.0,
    flip_tangent: torch.Tensor | None = None,
    phase_tangent: torch.Tensor | None = None,
    duration_tangent: torch.Tensor | None = None,
) -> tuple[torch.Tensor, torch.Tensor]:
    """Signal and its directional derivative along per-atom tissue tangents
    (names among :data:`CARRIED`: sum_k J_k t_k) and per-event tangents of the
    flips, phases and durations."""
    names = tuple(tissue_tangents)
    refuse_uncarried(names)
    plain = dataclasses.replace(tissue, buffers=[None if b is None else b.detach() for b in tissue.buffers])
    signal, jacobian = simulate_shaped(packed, plain, table, states=states, columns=names, factor=factor)
    tangent = torch.zeros_like(signal)
    for k, name in enumerate(names):
        tangent = tangent + jacobian[:, k] * tissue_tangents[name].detach().reshape(-1, 1).to(torch.float32)
    if any(d is not None for d in (flip_tangent, phase_tangent, duration_tangent)):
        tangent = tangent + event_tangent(packed, plain, table, states=states, factor=factor,
                                          flip_tangent=flip_tangent, phase_tangent=phase_tangent,
                                          duration_tangent=duration_tangent)
    return signal, tangent


def adjoint_semisolid(
    packed: PackedEvents,
    tissue: PreparedTissue,
    *,
    states: int,
    factor: complex,
    cotangent: torch.Tensor,
    flip: torch.Tensor | None = None,
    phase: torch.Tensor | None = None,
    duration: torch.Tensor | None = None,
    table: TransitionTable | None = None,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Gradients (events,) FP32 of a real loss with respect to every event's
    flip, phase and duration through the semisolid pool and/or slice-profiled
    pulses (``table``), given the loss's cotangent (voxels, outputs) on the
    signal ``simulate_shaped`` returns."""
    device = tissue.device
    _native.require_device(device)
    if tissue.exchanging:
        raise NotImplementedError("a semisolid pool beside an exchanging pool (three pools) is not built")
    lib = _native.load()
    lineshape = absorption(packed, tissue) if tissue.semisolid else None
    knots = _resident(lineshape, device) if lineshape is not None else None
    count = packed.event_count
    grads = [torch.empty(count, dtype=torch.float32, device=device) for _ in range(3)]
    events = _complex._events(packed, device, flip=flip, phase=phase, duration=duration)
    weight = cotangent.detach().to(device=device, dtype=torch.complex64)
    factor_re, factor_im = float(complex(factor).real), float(complex(factor).imag)
    shape_args = (None if knots is None else knots.data_ptr(), 0 if lineshape is None else lineshape.bins,
                  1.0 if lineshape is None else float(lineshape.step))
    with torch.cuda.device(device):
        if table is None:
            ptrs = _native.pointers([*(None if b is None else b.detach() for b in tissue.buffers), *events])
            weight = weight.contiguous()
            status = lib.tsim_semisolid_vjp(
                ptrs, tissue.atoms, count, states, packed.output_count, tissue.features, tissue.stride, factor_re,
                factor_im, *shape_args, weight.data_ptr(), *(g.data_ptr() for g in grads), _native.stream_of(device))
        else:
            # every voxel at every position; each position's cotangent the voxel's over their count
            locations = table.points
            spread = _across(tissue, locations)
            weight = (weight / locations).repeat_interleave(locations, dim=0).contiguous()
            profile = table.packed(device)
            ptrs = _native.pointers([*(None if b is None else b.detach() for b in spread.buffers), *events])
            status = lib.tsim_profiled_vjp(
                ptrs, spread.atoms, count, states, packed.output_count, spread.features, spread.stride, factor_re,
                factor_im, profile.data_ptr(), locations, table.bins, float(table.step), *shape_args,
                weight.data_ptr(), *(g.data_ptr() for g in grads), _native.stream_of(device))
    _native.check(status)
    return grads[0], grads[1], grads[2]


class SequenceSignal(torch.autograd.Function):
    """The signal of a tissue with a semisolid pool and/or slice-profiled
    pulses as a function of the events' flips, phases and durations and of
    the tissue maps it was differentiated along: ``backward`` takes the event
    gradients from the adjoint (``tsim_semisolid_vjp`` / ``tsim_profiled_vjp``)
    and the tissue gradients from the Jacobian columns the forward wrote (as
    ``ColumnSignal``) -- the reference's ``_NativeEpg.backward`` for these
    models (sequence/_accelerators.py:1544-1586). The signal and columns come
    in computed."""

    @staticmethod
    def forward(flip, phase, duration, signal, jacobian, run, *buffers):
        return signal.clone()

    @staticmethod
    def setup_context(ctx, inputs, output):
        flip, phase, duration, _, jacobian, run, *buffers = inputs
        ctx.save_for_backward(flip, phase, duration, *([] if jacobian is None else [jacobian]))
        ctx.save_for_forward(flip, phase, duration, *([] if jacobian is None else [jacobian]))
        ctx.run = run
        ctx.shapes = [None if b is None else b.shape for b in buffers]

    @staticmethod
    def jvp(ctx, d_flip, d_phase, d_duration, _d_signal, _d_jacobian, _run, *d_buffers):
        run = ctx.run
        flip, phase, duration, *held = ctx.saved_tensors
        tangent = event_tangent(run.packed, run.tissue, run.table, states=run.states, factor=run.factor,
                                flip_tangent=d_flip, phase_tangent=d_phase, duration_tangent=d_duration,
                                flip=flip, phase=phase, duration=duration)
        for k, d in enumerate(d_buffers):
            if d is not None:
                tangent = tangent + held[0][:, k] * d.reshape(-1, 1).to(torch.float32)
        return tangent

    @staticmethod
    def backward(ctx, grad):
        run = ctx.run
        flip, phase, duration, *held = ctx.saved_tensors
        needs = ctx.needs_input_grad
        out: list = [None, None, None, None, None, None]
        if any(needs[:3]):
            g_flip, g_phase, g_duration = adjoint_semisolid(
                run.packed, run.tissue, states=run.states, factor=run.factor, cotangent=grad, flip=flip, phase=phase,
                duration=duration, table=run.table)
            out[0] = g_flip.to(flip.device, flip.dtype) if needs[0] else None
            out[1] = g_phase.to(phase.device, phase.dtype) if needs[1] else None
            out[2] = g_duration.to(duration.device, duration.dtype) if needs[2] else None
        grads = []
        for k, (shape, need) in enumerate(zip(ctx.shapes, needs[6:])):
            if not need or shape is None:
                grads.append(None)
                continue
            g = (grad * held[0][:, k].conj()).real.sum(-1)
            grads.append(g.reshape(shape) if g.numel() == shape.numel() else g.sum().reshape(shape))
        return (*out, *grads)


def sequence_signal(
    packed: PackedEvents,
    tissue: PreparedTissue,
    table: TransitionTable | None,
    *,
    states: int,
    carried: tuple[str, ...],
    factor: complex = 1.0,
) -> torch.Tensor:
    """The signal of a tissue with a semisolid pool and/or slice-profiled
    pulses (``table``) with reverse-mode gradients reaching the packed flips,
    phases and durations (and the tissue maps ``carried`` names, among
    :data:`CARRIED`)."""
    from types import SimpleNamespace

    refuse_uncarried(carried)
    device = tissue.device
    plain = dataclasses.replace(tissue, buffers=[None if b is None else b.detach() for b in tissue.buffers])
    flip, phase, duration = (packed.flip.to(device), packed.phase.to(device), packed.duration.to(device))
    signal, jacobian = simulate_shaped(packed, plain, table, states=states, columns=carried, factor=factor,
                                       flip=flip, phase=phase, duration=duration)
    run = SimpleNamespace(packed=packed, tissue=plain, states=states, factor=factor, table=table)
    buffers = [tissue.buffers[_native.TISSUE_INDEX[name]] for name in carried]
    return SequenceSignal.apply(flip, phase, duration, signal, jacobian, run, *buffers)
