"""Launch of the complex EPG kernel (single pool or one exchanging pool,
diffusion, off-resonance, RF phase) with fused Jacobian columns."""

from __future__ import annotations

import ctypes

import torch

from .. import _native
from ._engine import PreparedTissue
from ._events import PackedEvents

_COLUMN_INDEX = {"t1_ms": 0, "t2_ms": 1, "m0": 2, "b1": 3}


def simulate_complex(
    packed: PackedEvents,
    tissue: PreparedTissue,
    *,
    states: int,
    columns: tuple[str, ...] = (),
    factor: complex = 1.0,
) -> tuple[torch.Tensor, torch.Tensor | None]:
    """Complex64 signal (atoms, outputs) and Jacobian (atoms, P, outputs)."""
    device = tissue.device
    _native.require_device(device)
    for name in columns:
        if name not in _COLUMN_INDEX:
            raise NotImplementedError(f"the complex kernel carries no tangent along {name!r}")
    lib = _native.load()
    events = list(packed.on(device))
    atoms, outputs = tissue.atoms, packed.output_count
    signal = torch.empty((atoms, outputs), dtype=torch.complex64, device=device)
    jacobian = torch.empty((atoms, len(columns), outputs), dtype=torch.complex64, device=device) if columns else None
    if atoms == 0 or outputs == 0:
        return signal, jacobian
    ptrs = _native.pointers([*tissue.buffers, *events])
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
