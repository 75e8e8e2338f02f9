"""Launch of the complex EPG kernel (single pool or one exchanging pool,
diffusion, off-resonance, RF phase) with fused Jacobian columns, and its
checkpointed adjoint behind an autograd Function."""

from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from .. import _native
from ._engine import PreparedTissue
from ._events import PackedEvents

_COLUMN_INDEX = {"t1_ms": 0, "t2_ms": 1, "m0": 2, "b1": 3}

#: Tissue fields the complex adjoint returns a per-atom gradient for.
GRADIENT_FIELDS = ("t1_ms", "t2_ms", "m0", "b1", "b1_phase_rad", "b0_hz", "inversion_efficiency",
                   "diffusion_um2_per_ms")


def _events(packed: PackedEvents, device: torch.device, flip: torch.Tensor | None = None,
            phase: torch.Tensor | None = None, duration: torch.Tensor | None = None) -> list[torch.Tensor]:
    """The nine event buffers on ``device``, with flip / phase / duration
    taken from the given tensors where they are given (detached: the kernels
    read values, autograd is the Function's business)."""
    events = list(packed.on(device))
    for slot, value in ((0, duration), (2, flip), (3, phase)):
        if value is not None:
            events[slot] = value.detach().to(device=device, dtype=torch.float32).contiguous()
    return [value.detach() for value in events]


def simulate_complex(
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
    """Complex64 signal (atoms, outputs) and Jacobian (atoms, P, outputs)."""
    device = tissue.device
    _native.require_device(device)
    for name in columns:
        if name not in _COLUMN_INDEX:
            raise NotImplementedError(f"the complex kernel carries no tangent along {name!r}")
    lib = _native.load()
    events = _events(packed, device, flip, phase, duration)
    atoms, outputs = tissue.atoms, packed.output_count
    signal = torch.empty((atoms, outputs), dtype=torch.complex64, device=device)
    jacobian = torch.empty((atoms, len(columns), outputs), dtype=torch.complex64, device=device) if columns else None
    if atoms == 0 or outputs == 0:
        return signal, jacobian
    ptrs = _native.pointers([*(None if b is None else b.detach() for b in tissue.buffers), *events])
    cols = (ctypes.c_int32 * max(1, len(columns)))(*[_COLUMN_INDEX[name] for name in columns])
    with torch.cuda.device(device):
        status = lib.tsim_complex_jacobian(
            ptrs,
            atoms,
            packed.event_count,
            states,
            outputs,
            2 if tissue.exchanging else 0,
            tissue.features,
            tissue.stride,
            cols,
            len(columns),
            float(complex(factor).real),
            float(complex(factor).imag),
            signal.data_ptr(),
            None if jacobian is None else jacobian.data_ptr(),
            _native.stream_of(device),
        )
    _native.check(status)
    return signal, jacobian


def adjoint_complex(
    packed: PackedEvents,
    tissue: PreparedTissue,
    *,
    states: int,
    factor: complex,
    cotangent: torch.Tensor,
    flip: torch.Tensor | None = None,
    phase: torch.Tensor | None = None,
    duration: torch.Tensor | None = None,
    segment_events: int = 0,
    tissue_wanted: tuple[str, ...] = (),
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, dict[str, torch.Tensor]]:
    """d(loss)/d(flip), d(loss)/d(phase) and d(loss)/d(duration), one entry
    per event, and d(loss)/d(field) per atom for each field of
    ``tissue_wanted`` (among GRADIENT_FIELDS), from the loss's complex
    cotangent on the signal (atoms, outputs) -- d(loss)/d(Re) + i
    d(loss)/d(Im), torch's convention. The state is checkpointed every
    ``segment_events`` events (0: about the square root of the event count)."""
    device = tissue.device
    _native.require_device(device)
    lib = _native.load()
    events = _events(packed, device, flip, phase, duration)
    count = packed.event_count
    grad_flip = torch.empty(count, dtype=torch.float32, device=device)
    grad_phase = torch.empty(count, dtype=torch.float32, device=device)
    grad_duration = torch.empty(count, dtype=torch.float32, device=device)
    grads = {}
    slots = (ctypes.c_void_p * _native.TISSUE_COUNT)()
    for name in tissue_wanted:
        if name not in GRADIENT_FIELDS:
            raise NotImplementedError(f"the complex adjoint returns no gradient for {name!r}")
        grads[name] = torch.empty(tissue.atoms, dtype=torch.float32, device=device)
        slots[_native.TISSUE_INDEX[name]] = grads[name].data_ptr()
    ptrs = _native.pointers([*(None if b is None else b.detach() for b in tissue.buffers), *events])
    w = cotangent.detach().to(device=device, dtype=torch.complex64).contiguous()
    with torch.cuda.device(device):
        status = lib.tsim_complex_vjp(
            ptrs,
            tissue.atoms,
            count,
            states,
            packed.output_count,
            2 if tissue.exchanging else 0,
            tissue.features,
            tissue.stride,
            float(complex(factor).real),
            float(complex(factor).imag),
            w.data_ptr(),
            int(segment_events),
            grad_flip.data_ptr(),
            grad_phase.data_ptr(),
            grad_duration.data_ptr(),
            slots,
            _native.TISSUE_COUNT if grads else 0,
            _native.stream_of(device),
        )
    _native.check(status)
    return grad_flip, grad_phase, grad_duration, grads


def adjoint_complex_jacobian(
    packed: PackedEvents,
    tissue: PreparedTissue,
    *,
    states: int,
    factor: complex,
    columns: tuple[str, ...],
    signal_cotangent: torch.Tensor | None,
    jacobian_cotangent: torch.Tensor | None,
    flip: torch.Tensor | None = None,
    phase: torch.Tensor | None = None,
    duration: torch.Tensor | None = None,
    segment_events: int = 0,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """d(loss)/d(flip), d(phase), d(duration) per event for a loss that reads
    the complex kernel's signal (atoms, outputs) and Jacobian (atoms, P,
    outputs) along ``columns`` (among t1_ms, t2_ms, m0, b1), from their
    cotangents (torch's convention): one first-order pass for the signal and
    one forward-over-reverse pass per T1, T2 or B1 column."""
    device = tissue.device
    _native.require_device(device)
    for name in columns:
        if name not in _COLUMN_INDEX:
            raise NotImplementedError(f"the complex kernel carries no tangent along {name!r}")
    lib = _native.load()
    events = _events(packed, device, flip, phase, duration)
    count = packed.event_count
    grads = [torch.empty(count, dtype=torch.float32, device=device) for _ in range(3)]
    ptrs = _native.pointers([*(None if b is None else b.detach() for b in tissue.buffers), *events])
    cols = (ctypes.c_int32 * max(1, len(columns)))(*[_COLUMN_INDEX[name] for name in columns])

    def held(value):
        return None if value is None else value.detach().to(device=device, dtype=torch.complex64).contiguous()

    ws, wj = held(signal_cotangent), held(jacobian_cotangent)
    with torch.cuda.device(device):
        status = lib.tsim_complex_vjp_jacobian(
            ptrs, tissue.atoms, count, states, packed.output_count, 2 if tissue.exchanging else 0,
            tissue.features, tissue.stride, float(complex(factor).real), float(complex(factor).imag),
            None if ws is None else ws.data_ptr(), cols, len(columns), None if wj is None else wj.data_ptr(),
            int(segment_events), grads[0].data_ptr(), grads[1].data_ptr(), grads[2].data_ptr(),
            _native.stream_of(device),
        )
    _native.check(status)
    return grads[0], grads[1], grads[2]


class ComplexJacobian(torch.autograd.Function):
    """The complex kernel's signal and Jacobian as a function of the events'
    flips, phases and durations, with the reference's forward-over-reverse
    backward (``_NativeEpgJvp.backward``, sequence/_accelerators.py:
    1787-1850; ``simulate_vjp_jvp``, _epg_cpu.cpp:7759-9013): what a design
    loss on a phase-cycled, off-resonant, exchanging or diffusing protocol's
    Jacobian needs. Tissue maps are constants here."""

    @staticmethod
    def forward(flip, phase, duration, run, columns):
        return simulate_complex(run.packed, run.tissue, states=run.states, columns=columns, factor=run.factor,
                                flip=flip, phase=phase, duration=duration)

    @staticmethod
    def setup_context(ctx, inputs, output):
        flip, phase, duration, run, columns = inputs
        ctx.save_for_backward(flip, phase, duration)
        ctx.run = run
        ctx.columns = columns

    @staticmethod
    def backward(ctx, grad_signal, grad_jacobian):
        flip, phase, duration = ctx.saved_tensors
        run = ctx.run
        g_flip, g_phase, g_duration = adjoint_complex_jacobian(
            run.packed, run.tissue, states=run.states, factor=run.factor, columns=ctx.columns,
            signal_cotangent=grad_signal, jacobian_cotangent=grad_jacobian if ctx.columns else None,
            flip=flip, phase=phase, duration=duration, segment_events=run.checkpoint_events)
        needs = ctx.needs_input_grad
        return (g_flip.to(flip.dtype) if needs[0] else None, g_phase.to(phase.dtype) if needs[1] else None,
                g_duration.to(duration.dtype) if needs[2] else None, None, None)


def jvp_complex(
    packed: PackedEvents,
    tissue: PreparedTissue,
    *,
    states: int,
    factor: complex,
    tissue_tangents: dict[str, torch.Tensor],
    flip: torch.Tensor | None = None,
    phase: torch.Tensor | None = None,
    duration: torch.Tensor | None = None,
    flip_tangent: torch.Tensor | None = None,
    phase_tangent: torch.Tensor | None = None,
    duration_tangent: torch.Tensor | None = None,
    with_signal: bool = False,
) -> torch.Tensor | tuple[torch.Tensor, torch.Tensor]:
    """The directional derivative (atoms, outputs) complex64 of the signal
    ``factor * sample`` along a tangent on tissue fields (``tissue_tangents``:
    field name -> a tensor laid out like that field's prepared buffer) and on
    the events' flips, phases and durations ((events,) each, or None); with
    ``with_signal``, (signal, tangent) from the same launch."""
    device = tissue.device
    _native.require_device(device)
    lib = _native.load()
    events = _events(packed, device, flip, phase, duration)
    atoms, outputs = tissue.atoms, packed.output_count
    tangent = torch.empty((atoms, outputs), dtype=torch.complex64, device=device)
    signal = torch.empty_like(tangent) if with_signal else None
    if atoms == 0 or outputs == 0:
        return (signal, tangent) if with_signal else tangent
    held = []
    slots = (ctypes.c_void_p * (_native.TISSUE_COUNT + 3))()
    for name, value in tissue_tangents.items():
        if name not in GRADIENT_FIELDS:
            raise NotImplementedError(f"the complex JVP carries no tangent along {name!r}")
        buffer = tissue.buffers[_native.TISSUE_INDEX[name]]
        flat = value.detach().to(device=device, dtype=torch.float32).reshape(-1)
        if buffer is not None and flat.numel() != buffer.numel():
            flat = flat.expand(buffer.numel())
        flat = flat.contiguous()
        held.append(flat)
        slots[_native.TISSUE_INDEX[name]] = flat.data_ptr()
    for offset, value in enumerate((duration_tangent, flip_tangent, phase_tangent)):
        if value is not None:
            flat = value.detach().to(device=device, dtype=torch.float32).reshape(-1).contiguous()
            held.append(flat)
            slots[_native.TISSUE_COUNT + offset] = flat.data_ptr()
    ptrs = _native.pointers([*(None if b is None else b.detach() for b in tissue.buffers), *events])
    with torch.cuda.device(device):
        status = lib.tsim_complex_jvp(
            ptrs, atoms, packed.event_count, states, outputs, 2 if tissue.exchanging else 0, tissue.features,
            tissue.stride, float(complex(factor).real), float(complex(factor).imag), slots,
            _native.TISSUE_COUNT + 3, None if signal is None else signal.data_ptr(), tangent.data_ptr(),
            _native.stream_of(device),
        )
    _native.check(status)
    return (signal, tangent) if with_signal else tangent


@dataclasses.dataclass
class SignalRun:
    """What a signal evaluation fixes besides its tensors: the protocol, the
    tissue's layout (its buffers are the Function's inputs), the state count,
    the output factor, the real-subspace verdict with its flip and sample
    signs, and the checkpoint spacings of the two adjoints."""

    packed: PackedEvents
    tissue: PreparedTissue
    states: int
    factor: complex
    real: bool
    flip_signs: torch.Tensor | None
    samples: torch.Tensor | None
    checkpoint_pulses: int = 64
    checkpoint_events: int = 0

    def with_buffers(self, buffers) -> PreparedTissue:
        return dataclasses.replace(self.tissue, buffers=list(buffers))


class EpgSignal(torch.autograd.Function):
    """The recorded signal (atoms, outputs), complex64, as a function of the
    events' flips, phases and durations and of the 17 prepared tissue
    buffers: the reference's ``_NativeEpg`` (sequence/_accelerators.py:
    1402-1594) -- forward on the real-subspace kernel where the verdict
    allows it, else on the complex kernel; ``backward`` through the
    real-subspace adjoint when what requires grad is among what it carries
    (flips, durations, T1, T2, M0, B1, inversion efficiency), else through
    the complex adjoint (adding phases, B1phase, B0, diffusion); ``jvp``
    along any direction of those inputs through the complex JVP kernel
    (reference ``_NativeEpg.jvp``, :1511-1542), under ``torch.func.jvp`` and
    ``torch.autograd.forward_ad`` alike."""

    @staticmethod
    def forward(flip, phase, duration, *rest):
        *buffers, run = rest
        tissue = run.with_buffers(buffers)
        if run.real:
            signed = flip if run.flip_signs is None else flip * run.flip_signs.to(flip.device)
            from . import _engine

            signal, _ = _engine.simulate_real(run.packed, tissue, states=run.states, factor=run.factor,
                                              flip=signed, samples=run.samples, duration=duration)
            return signal
        signal, _ = simulate_complex(run.packed, tissue, states=run.states, factor=run.factor, flip=flip,
                                     phase=phase, duration=duration)
        return signal

    @staticmethod
    def setup_context(ctx, inputs, output):
        flip, phase, duration, *rest = inputs
        *buffers, run = rest
        ctx.save_for_backward(flip, phase, duration, *buffers)
        ctx.save_for_forward(flip, phase, duration, *buffers)
        ctx.run = run

    @staticmethod
    def backward(ctx, grad_signal):
        from . import _engine

        run = ctx.run
        flip, phase, duration, *buffers = ctx.saved_tensors
        tissue = run.with_buffers(buffers)
        needs = ctx.needs_input_grad
        wanted = tuple(name for name, need in zip(_native.TISSUE_NAMES, needs[3:3 + _native.TISSUE_COUNT]) if need)
        out_tissue: list = [None] * _native.TISSUE_COUNT
        shifting = run.states < 2 and bool(((run.packed.action.cpu().numpy().astype(np.int32) & 19) != 0).any())
        if run.real and not needs[1] and not shifting and all(name in _engine.GRADIENT_FIELDS for name in wanted):
            # every output is c v, v real, c = factor i (sample sign): the
            # loss's derivative along v is Re(conj(c) grad)
            scale = torch.tensor(complex(run.factor) * 1j, dtype=torch.complex64, device=grad_signal.device).conj()
            if run.samples is not None:
                scale = scale * run.samples.to(device=grad_signal.device, dtype=torch.complex64)
            signs = None if run.flip_signs is None else run.flip_signs.to(flip.device)
            signed = flip if signs is None else flip * signs
            zero = _engine.zero_intervals_reached(run.packed) if needs[2] else None
            g_flip, g_duration, grads = _engine.adjoint_real(
                run.packed, tissue, states=run.states, columns=(), flip=signed, duration=duration,
                signal_cotangent=(grad_signal * scale).real, jacobian_cotangent=None,
                checkpoint_pulses=run.checkpoint_pulses, tissue_wanted=wanted, zero_intervals=zero)
            if signs is not None:
                g_flip = g_flip * signs
            g_phase = None
        else:
            g_flip, g_phase, g_duration, grads = adjoint_complex(
                run.packed, tissue, states=run.states, factor=run.factor, cotangent=grad_signal, flip=flip,
                phase=phase, duration=duration, segment_events=run.checkpoint_events, tissue_wanted=wanted)
        for name, g in grads.items():
            index = _native.TISSUE_INDEX[name]
            buffer = buffers[index]
            if g is not None and buffer is not None and g.numel() != buffer.numel():
                g = g.sum().reshape(buffer.shape)  # one value for every atom
            out_tissue[index] = g
        return (
            g_flip.to(flip.dtype) if needs[0] else None,
            g_phase.to(phase.dtype) if needs[1] and g_phase is not None else None,
            g_duration.to(duration.dtype) if needs[2] else None,
            *out_tissue,
            None,
        )

    @staticmethod
    def jvp(ctx, d_flip, d_phase, d_duration, *rest):
        *d_buffers, _ = rest
        flip, phase, duration, *buffers = ctx.saved_tensors
        return _EpgTangent.apply(flip, phase, duration, d_flip, d_phase, d_duration, ctx.run, *buffers, *d_buffers)


class _EpgTangent(torch.autograd.Function):
    """The complex JVP kernel behind ``EpgSignal.jvp``: a Function of its
    own so that its forward sees plain tensors whichever transform asked."""

    @staticmethod
    def forward(flip, phase, duration, d_flip, d_phase, d_duration, run, *both):
        count = _native.TISSUE_COUNT
        buffers, d_buffers = both[:count], both[count:]
        tissue = run.with_buffers(buffers)
        tangents = {}
        for name, value in zip(_native.TISSUE_NAMES, d_buffers):
            if value is None:
                continue
            if name not in GRADIENT_FIELDS:
                raise NotImplementedError(f"forward mode carries no tangent along {name!r}")
            tangents[name] = value
        return jvp_complex(run.packed, tissue, states=run.states, factor=run.factor, tissue_tangents=tangents,
                           flip=flip, phase=phase, duration=duration, flip_tangent=d_flip,
                           phase_tangent=d_phase, duration_tangent=d_duration)

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, grad):
        raise NotImplementedError("a directional derivative is not differentiated again (forward-over-reverse "
                                  "runs through Simulator.jacobian with reverse mode)")
