"""Launch of the semisolid-pool (magnetization-transfer) kernel.

Reference: the SEMISOLID pool kind of the complex state machine
(src/torchsim/_epg_cpu.cpp:4216-4566 with pool kind 1) as
``simulate_native`` reaches it with an absorption table sized by
``_absorption_table`` (sequence/_simulation.py:80-94, 466-470).
"""

from __future__ import annotations

import ctypes
import threading

import torch

from .. import _native
from ._engine import PreparedTissue
from ._events import PackedEvents
from ._lineshape import Lineshape, lineshape_reaching

_COLUMN_INDEX = {"t1_ms": 0, "t2_ms": 1, "m0": 2, "b1": 3}
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
) -> tuple[torch.Tensor, torch.Tensor | None]:
    """Complex64 signal (atoms, outputs) and Jacobian (atoms, P, outputs) of a
    tissue with a semisolid pool; columns among t1_ms, t2_ms, m0, b1."""
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
    events = [b.detach() for b in packed.on(device)]
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
