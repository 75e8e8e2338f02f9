"""The C-ABI library loads and exports every symbol its header declares; error
paths return statuses and messages without touching a GPU."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> list[str]:
    text = (ROOT / "include" / "torchsim_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsim_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    for required in ("tsim_simulate", "tsim_simulate_jacobian", "tsim_last_error", "tsim_complex_simulate"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from torchsim_amd import _native

    if not _native.LIBRARY.exists():
        pytest.skip("library not built")
    lib = ctypes.CDLL(str(_native.LIBRARY))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    # Every declared symbol has a ctypes signature on the Python side too.
    assert set(declared_symbols()) <= set(_native.SIGNATURES)


def test_bad_arguments_are_statuses_not_crashes():
    from torchsim_amd import _native

    if not _native.LIBRARY.exists():
        pytest.skip("library not built")
    lib = _native.load()
    ptrs = (ctypes.c_void_p * 3)()
    status = lib.tsim_simulate(ptrs, 3, 1, 1, 1, 10, 1, 1, 1, 0, 0, 1, None)
    assert status == 1
    assert b"wrong number of buffer pointers" in lib.tsim_last_error()
    cols = (ctypes.c_int32 * 1)(5)
    packed = (ctypes.c_void_p * 26)()
    status = lib.tsim_simulate_jacobian(packed, 1, 1, 0, 1, 0, 1, cols, 1, 1.0, 0.0, None, None, None, -1, None)
    assert status == 1  # zero states
    with pytest.raises(ValueError):
        _native.check(1)
    assert lib.tsim_tile_count() >= 10
    shapes = _native.tiles()
    assert all(sl * lanes <= 64 for sl, lanes in shapes)


def test_projection_arguments_are_checked_on_the_host():
    """tsim_simulate_projected refuses a rank outside 1..16, a missing basis
    and diffusion before it touches the device."""
    from torchsim_amd import _native

    if not _native.LIBRARY.exists():
        pytest.skip("library not built")
    lib = _native.load()
    packed = (ctypes.c_void_p * 26)()
    cols = (ctypes.c_int32 * 1)(0)
    for rank in (0, 17):
        status = lib.tsim_simulate_projected(packed, 4, 1, 10, 1, 0, 1, cols, 1, 0.0, 1.0, None, 1, rank, None,
                                             None, -1, None)
        assert status == 2 and b"basis directions" in lib.tsim_last_error()
    status = lib.tsim_simulate_projected(packed, 4, 1, 10, 1, 0, 1, cols, 1, 0.0, 1.0, None, None, 8, None, None,
                                         -1, None)
    assert status == 1 and b"basis buffer is null" in lib.tsim_last_error()
    status = lib.tsim_simulate_projected(packed, 4, 1, 10, 1, 0, 1, cols, 1, 0.0, 1.0, None, 1, 8, None, None, -1,
                                         None)
    assert status == 1 and b"Jacobian buffer is null" in lib.tsim_last_error()
    status = lib.tsim_simulate_projected(packed, 4, 1, 10, 1, _native.FEATURE_DIFFUSING, 1, cols, 0, 0.0, 1.0, None,
                                         1, 8, None, None, -1, None)
    assert status == 2 and b"diffusion" in lib.tsim_last_error()


def test_noalloc_entry_point_rejects_streams_beyond_its_workspace():
    """The allocation-free entry point holds at most tsim_noalloc_max_events()
    events; a longer stream is a status 1 with a message, checked on the host
    before anything reaches a device (the pointers here are never read)."""
    from torchsim_amd import _native

    if not _native.LIBRARY.exists():
        pytest.skip("library not built")
    lib = _native.load()
    limit = lib.tsim_noalloc_max_events()
    assert limit == 16384
    packed = (ctypes.c_void_p * 26)()
    packed[0] = packed[1] = 16  # t1, t2: never dereferenced on this path
    cols = (ctypes.c_int32 * 1)(0)
    status = lib.tsim_simulate_jacobian_noalloc(packed, 8, limit + 1, 10, 4, 0, 1, cols, 1, 0.0, 1.0, None,
                                                16, 16, -1, None)
    assert status == 1
    assert b"at most 16384 events" in lib.tsim_last_error()
