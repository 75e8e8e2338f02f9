"""Launch of the complex EPG kernel (single pool or one exchanging pool,
diffusion, off-resonance, RF phase) with fused Jacobian columns, and its
checkpointed adjoint behind an autograd Function."""

from __future__ import annotations

import ctypes

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


class FusedComplexSignal(torch.autograd.Function):
    """The complex kernel's signal, differentiable in the event flips, phases
    and durations and in the per-atom fields of GRADIENT_FIELDS (reference
    sequence/_accelerators.py:1402-1586: ``_NativeEpg``'s forward and its
    first-order backward through ``simulate_vjp``, here one forward kernel and
    one checkpointed adjoint kernel). The tissue inputs are the prepared
    per-atom buffers the kernels read, so autograd carries their gradients
    back to the caller's maps."""

    @staticmethod
    def forward(ctx, flip, phase, duration, t1, t2, m0, b1, b1_phase, b0, ie, diffusion, packed, tissue, states,
                factor, segment_events):
        signal, _ = simulate_complex(packed, tissue, states=states, factor=factor, flip=flip, phase=phase,
                                     duration=duration)
        ctx.save_for_backward(flip, phase, duration)
        ctx.shapes = [None if v is None else v.shape for v in (t1, t2, m0, b1, b1_phase, b0, ie, diffusion)]
        ctx.packed = packed
        ctx.tissue = tissue
        ctx.states = states
        ctx.factor = complex(factor)
        ctx.segment_events = segment_events
        return signal

    @staticmethod
    def backward(ctx, grad_signal):
        flip, phase, duration = ctx.saved_tensors
        wanted = tuple(name for name, need in zip(GRADIENT_FIELDS, ctx.needs_input_grad[3:11]) if need)
        grad_flip, grad_phase, grad_duration, grads = adjoint_complex(
            ctx.packed, ctx.tissue, states=ctx.states, factor=ctx.factor, cotangent=grad_signal, flip=flip,
            phase=phase, duration=duration, segment_events=ctx.segment_events, tissue_wanted=wanted,
        )
        tissue_out = []
        for name, shape in zip(GRADIENT_FIELDS, ctx.shapes):
            g = grads.get(name)
            if g is not None and shape is not None and g.numel() != torch.Size(shape).numel():
                g = g.sum().reshape(shape)  # one value for every atom
            tissue_out.append(g)
        return (
            grad_flip.to(flip.dtype) if ctx.needs_input_grad[0] else None,
            grad_phase.to(phase.dtype) if ctx.needs_input_grad[1] else None,
            grad_duration.to(duration.dtype) if ctx.needs_input_grad[2] else None,
            *tissue_out,
            None, None, None, None, None,
        )
