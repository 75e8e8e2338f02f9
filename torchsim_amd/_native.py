"""The C ABI of ``libtorchsim_hip.so``, loaded through ctypes.

The library is the only compute path: there is no CPU fallback. If it is
missing, or no GPU is visible, every simulation raises rather than answering
from somewhere else.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Any

import torch

LIBRARY = Path(__file__).resolve().parent / "libtorchsim_hip.so"

#: Tissue property order of the packed ABI (reference sequence/_parameters.py:105-133).
TISSUE_NAMES = (
    "t1_ms",
    "t2_ms",
    "m0",
    "b1",
    "b1_phase_rad",
    "b0_hz",
    "inversion_efficiency",
    "diffusion_um2_per_ms",
    "velocity_m_per_s",
    "bound_fraction",
    "bound_exchange_rate",
    "t1_bound_ms",
    "pool_b_fraction",
    "pool_b_exchange_rate",
    "t1_pool_b_ms",
    "t2_pool_b_ms",
    "pool_b_shift_hz",
)
TISSUE_INDEX = {name: index for index, name in enumerate(TISSUE_NAMES)}
TISSUE_COUNT = len(TISSUE_NAMES)
EVENT_COUNT = 9
PACKED_COUNT = TISSUE_COUNT + EVENT_COUNT

# Feature bits (reference src/torchsim/_epg_cpu.cpp:89-96).
FEATURE_OFF_AXIS = 1 << 0
FEATURE_MOVING = 1 << 1
FEATURE_DIFFUSING = 1 << 2
FEATURE_TRANSMIT = 1 << 3
FEATURE_DENSITY = 1 << 4
FEATURE_INVERTING = 1 << 5

_STATUS = {1: ValueError, 2: NotImplementedError, 3: RuntimeError}

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int
_F32 = ctypes.c_float

#: Every symbol include/torchsim_hip.h declares, with its ctypes signature.
#: ABI revision these bindings are written against (``tsim_version``).
ABI_VERSION = 3

SIGNATURES: dict[str, tuple[Any, list[Any]]] = {
    "tsim_last_error": (ctypes.c_char_p, []),
    "tsim_version": (_I32, []),
    "tsim_simulate": (
        _I32,
        [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I64, _I64, _I32, _I64, _I32, _I32, _I64, _P],
    ),
    "tsim_complex_simulate": (
        _I32,
        [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I64, _I32, _I32, _I64, _P],
    ),
    "tsim_simulate_jacobian": (
        _I32,
        [
            ctypes.POINTER(_P),
            _I64,
            _I64,
            _I64,
            _I64,
            _I32,
            _I64,
            ctypes.POINTER(ctypes.c_int32),
            _I64,
            _F32,
            _F32,
            _P,
            _P,
            _P,
            _I32,
            _P,
        ],
    ),
    "tsim_simulate_jacobian_noalloc": (
        _I32,
        [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I64, ctypes.POINTER(ctypes.c_int32), _I64, _F32, _F32,
         _P, _P, _P, _I32, _P],
    ),
    "tsim_noalloc_max_events": (_I64, []),
    "tsim_simulate_projected": (
        _I32,
        [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I64, ctypes.POINTER(ctypes.c_int32), _I64, _F32, _F32,
         _P, _P, _I64, _P, _P, _I32, _P],
    ),
    "tsim_complex_jacobian": (
        _I32,
        [
            ctypes.POINTER(_P),
            _I64,
            _I64,
            _I64,
            _I64,
            _I32,
            _I32,
            _I64,
            ctypes.POINTER(ctypes.c_int32),
            _I64,
            _F32,
            _F32,
            _P,
            _P,
            _P,
        ],
    ),
    "tsim_simulate_vjp_jacobian": (
        _I32,
        [
            ctypes.POINTER(_P),
            _I64,
            _I64,
            _I64,
            _I64,
            _I32,
            _I64,
            ctypes.POINTER(ctypes.c_int32),
            _I64,
            _P,
            _P,
            _I64,
            _I64,
            _P,
            _P,
            ctypes.POINTER(_P),
            _I64,
            _P,
            _P,
        ],
    ),
    "tsim_gauss_newton_normal": (
        _I32,
        [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I64, ctypes.POINTER(ctypes.c_int32), _I64,
         _F32, _F32, _P, _P, _P, _I32, _P],
    ),
    "tsim_normal_from_jacobian": (_I32, [_P, _P, _P, _I64, _I64, _I32, _P, _P]),
    "tsim_lm_propose": (_I32, [_P, _P, _P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double), _I64, _I32, _P, _P, _P, _P, _P]),
    "tsim_lm_accept": (_I32, [_P, _P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P, _P]),
    "tsim_match_operand_size": (_I32, [_I64, _I64, _I32, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]),
    "tsim_match_planes": (_I32, [_P, _I64, _I64, _P, _P]),
    "tsim_match_layout_atoms": (_I32, [_P, _I32, _I64, _I64, _I32, _P, _P, _P, _P]),
    "tsim_match": (_I32, [_P, _P, _I32, _I64, _P, _I32, _P, _P, _I32, _I64, _I64, _I32, _I64, _P, _P, _P, _P, _P]),
    "tsim_solve_damped": (_I32, [_P, _P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P]),
    "tsim_normal_stride": (_I32, []),
    "tsim_tile_count": (_I32, []),
    "tsim_tile_shape": (_I32, [_I32, ctypes.POINTER(_I32), ctypes.POINTER(_I32)]),
}

class BatchRequest(ctypes.Structure):
    """Mirror of ``tsim_batch_request`` (include/torchsim_hip.h)."""

    _fields_ = [
        ("packed", _P * PACKED_COUNT),
        ("atom_count", _I64),
        ("event_count", _I64),
        ("state_count", _I64),
        ("output_count", _I64),
        ("tissue_stride", _I64),
        ("features", ctypes.c_int32),
        ("column_count", ctypes.c_int32),
        ("columns", ctypes.c_int32 * 5),
        ("factor_re", _F32),
        ("factor_im", _F32),
        ("sample_sign", _P),
        ("signal", _P),
        ("jacobian", _P),
        ("data", _P),
        ("normal", _P),
    ]


SIGNATURES["tsim_complex_vjp"] = (
    _I32,
    [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I32, _I64, _F32, _F32, _P, _I64, _P, _P, _P,
     ctypes.POINTER(_P), _I64, _P],
)
SIGNATURES["tsim_complex_vjp_jacobian"] = (
    _I32,
    [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I32, _I64, _F32, _F32, _P, ctypes.POINTER(ctypes.c_int32),
     _I64, _P, _I64, _P, _P, _P, _P],
)
SIGNATURES["tsim_semisolid_jacobian"] = (
    _I32,
    [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I64, ctypes.POINTER(ctypes.c_int32), _I64, _F32, _F32, _P,
     _I64, _F32, _P, _P, _P],
)
SIGNATURES["tsim_semisolid_vjp"] = (
    _I32,
    [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I64, _F32, _F32, _P, _I64, _F32, _P, _P, _P, _P, _P],
)
SIGNATURES["tsim_profiled_vjp"] = (
    _I32,
    [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I64, _F32, _F32, _P, _I64, _I64, _F32, _P, _I64, _F32, _P, _P,
     _P, _P, _P],
)
SIGNATURES["tsim_semisolid_event_jvp"] = (
    _I32,
    [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I64, _F32, _F32, _P, _I64, _I64, _F32, _P, _I64, _F32, _P, _P,
     _P, _P, _P],
)
SIGNATURES["tsim_profiled_jacobian"] = (
    _I32,
    [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I64, ctypes.POINTER(ctypes.c_int32), _I64, _F32, _F32, _P,
     _I64, _I64, _F32, _P, _I64, _F32, _P, _P, _P],
)
SIGNATURES["tsim_complex_jvp"] = (
    _I32,
    [ctypes.POINTER(_P), _I64, _I64, _I64, _I64, _I32, _I32, _I64, _F32, _F32, ctypes.POINTER(_P), _I64, _P, _P,
     _P],
)
SIGNATURES["tsim_simulate_batch"] = (_I32, [ctypes.POINTER(BatchRequest), _I64, _I32, _I32, _I64, _P])
SIGNATURES["tsim_batch_request_size"] = (_I32, [])
SIGNATURES["tsim_complex_batch"] = (_I32, [ctypes.POINTER(BatchRequest), _I64, _I64, _P])

_lib: ctypes.CDLL | None = None


def load(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load the library (once) and bind every declared signature.

    Raises
    ------
    RuntimeError
        If the library has not been built.
    """
    global _lib
    if _lib is not None and path is None:
        return _lib
    where = Path(path) if path is not None else LIBRARY
    if not where.exists():
        raise RuntimeError(
            f"{where} is missing: build it with `python -m torchsim_amd._build` "
            "(the HIP library is the only compute path; there is no fallback)"
        )
    lib = ctypes.CDLL(str(where))
    for name, (restype, argtypes) in SIGNATURES.items():
        function = getattr(lib, name)
        function.restype = restype
        function.argtypes = argtypes
    if lib.tsim_version() != ABI_VERSION:
        raise RuntimeError(f"{where} has ABI revision {lib.tsim_version()}, these bindings expect {ABI_VERSION}: "
                           "rebuild it with `python -m torchsim_amd._build`")
    if lib.tsim_batch_request_size() != ctypes.sizeof(BatchRequest):
        raise RuntimeError("tsim_batch_request layout differs between the library and the bindings")
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    """Raise the exception a nonzero status stands for."""
    if status == 0:
        return
    message = load().tsim_last_error().decode(errors="replace")
    raise _STATUS.get(status, RuntimeError)(message)


def require_device(device: torch.device) -> None:
    """Refuse anything but a visible HIP device."""
    if device.type != "cuda":
        raise RuntimeError(
            f"torchsim_amd computes on the GPU only; got tensors on {device}"
        )
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU is visible: torchsim_amd has no CPU path")


def stream_of(device: torch.device) -> int:
    """The raw hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


def pointers(tensors: list[torch.Tensor | None]) -> ctypes.Array:
    """A C array of device pointers (null for ``None``)."""
    array = (_P * len(tensors))()
    for index, tensor in enumerate(tensors):
        array[index] = None if tensor is None else tensor.data_ptr()
    return array


def tiles() -> list[tuple[int, int]]:
    """The (orders per lane, lanes per atom) of every compiled state tile."""
    lib = load()
    shapes = []
    for tile in range(lib.tsim_tile_count()):
        sl, lanes = _I32(), _I32()
        check(lib.tsim_tile_shape(tile, ctypes.byref(sl), ctypes.byref(lanes)))
        shapes.append((sl.value, lanes.value))
    return shapes
