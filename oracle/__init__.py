"""FP64 CPU oracle for the EPG hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker or the timed CPU
baseline. The product (``torchsim_amd``) never imports it.

* :func:`real` -- ctypes binding of ``epg_oracle.c``: the real-subspace state
  machine in double precision with forward-mode tangents (reference
  src/torchsim/_epg_cpu.cpp:3041-3285).
* :mod:`oracle.packed` -- the complex state machine in torch float64
  (reference tests/utils/packed_reference.py:36-389), differentiable by
  autograd, for the complex, exchanging-pool and diffusion paths and for
  sequence gradients.
* :mod:`oracle.isochromat` -- a Bloch isochromat brute force, independent of
  EPG, used to pin the oracle itself.
* :func:`reference_module` -- the reference's own CPU extension built from its
  sources into ``oracle/_ref`` (a second checker; CPU baseline "reference").

Parity status of each function is stated in its docstring.
"""

from __future__ import annotations

import ctypes
import importlib.machinery
import importlib.util
import subprocess
from pathlib import Path
from typing import Any

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = HERE / "libepg_oracle.so"
_lib = None


def build() -> None:
    """Compile the C oracle (and the reference extension where its sources exist)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if Path("/root/reference/src/torchsim/_epg_cpu.cpp").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def _load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        lib = ctypes.CDLL(str(_LIB))
        d = ctypes.POINTER(ctypes.c_double)
        i32 = ctypes.POINTER(ctypes.c_int32)
        u8 = ctypes.POINTER(ctypes.c_uint8)
        i64 = ctypes.c_int64
        lib.epg_oracle_real.restype = ctypes.c_int
        lib.epg_oracle_real.argtypes = [i64, i64, i64, i64, d, d, d, d, d, d, i32, d, i32, d, u8, i32, i64, d, d, d, d]
        _lib = lib
    return _lib


def _f64(value: Any, atoms: int | None = None) -> np.ndarray | None:
    if value is None:
        return None
    array = np.ascontiguousarray(np.asarray(value, dtype=np.float64).reshape(-1))
    return array


def _ptr(array: np.ndarray | None, ctype=ctypes.c_double):
    if array is None:
        return None
    return array.ctypes.data_as(ctypes.POINTER(ctype))


#: Tissue tangent rows of the C oracle.
TANGENT_ROWS = ("t1", "t2", "m0", "b1", "ie", "diffusion")


def real(
    events: dict[str, Any],
    *,
    t1: Any,
    t2: Any,
    m0: Any = None,
    b1: Any = None,
    ie: Any = None,
    diffusion: Any = None,
    states: int,
    outputs: int,
    tangents: list[dict[str, Any]] | None = None,
) -> tuple[np.ndarray, np.ndarray | None]:
    """Real-subspace signal ``m0 * a0`` (atoms, outputs) and its tangents.

    ``events`` holds numpy arrays ``duration, kind, flip, action,
    output_index`` (the packed stream). Each tangent direction is a dict with
    optional per-atom rows ``t1, t2, m0, b1, ie, diffusion`` and per-event rows
    ``duration, flip``. Returns ``(signal, tangents)`` with tangents shaped
    (atoms, directions, outputs).

    Parity: pinned -- benchmark checksums, the reference's own CPU kernel,
    isochromat brute force, finite differences, closed forms
    (tests/test_oracle.py).
    """
    lib = _load()
    t1 = _f64(t1)
    t2 = _f64(t2)
    atoms = int(max(t1.size, t2.size, *(np.size(v) for v in (m0, b1, ie, diffusion) if v is not None)))
    bufs = [t1, t2, _f64(m0), _f64(b1), _f64(ie), _f64(diffusion)]
    per_atom = np.array([0 if b is None else int(b.size == atoms and atoms > 1) for b in bufs], dtype=np.int32)
    for k, b in enumerate(bufs):
        if b is not None and b.size not in (1, atoms):
            raise ValueError(f"tissue row {k} has {b.size} entries for {atoms} atoms")
    duration = _f64(events["duration"])
    kind = np.ascontiguousarray(np.asarray(events["kind"], dtype=np.int32))
    flip = _f64(events["flip"])
    action = np.ascontiguousarray(np.asarray(events["action"], dtype=np.uint8))
    output_index = np.ascontiguousarray(np.asarray(events["output_index"], dtype=np.int32))
    count = kind.size
    directions = tangents or []
    ntan = len(directions)
    tan_tissue = tan_events = None
    if ntan:
        tan_tissue = np.zeros((ntan, len(TANGENT_ROWS), atoms))
        tan_events = np.zeros((ntan, 2, count))
        for k, direction in enumerate(directions):
            for q, name in enumerate(TANGENT_ROWS):
                if name in direction:
                    tan_tissue[k, q] = np.broadcast_to(np.asarray(direction[name], dtype=np.float64), (atoms,))
            for q, name in enumerate(("duration", "flip")):
                if name in direction:
                    tan_events[k, q] = np.asarray(direction[name], dtype=np.float64)
    signal = np.zeros((atoms, outputs))
    tangent_out = np.zeros((atoms, ntan, outputs)) if ntan else None
    status = lib.epg_oracle_real(
        atoms, count, states, outputs,
        *[_ptr(b) for b in bufs],
        _ptr(per_atom, ctypes.c_int32),
        _ptr(duration), _ptr(kind, ctypes.c_int32), _ptr(flip), _ptr(action, ctypes.c_uint8),
        _ptr(output_index, ctypes.c_int32),
        ntan, _ptr(tan_tissue), _ptr(tan_events),
        _ptr(signal), _ptr(tangent_out),
    )
    if status != 0:
        raise RuntimeError(f"oracle failed with status {status}")
    return signal, tangent_out


def events_of(packed: Any) -> dict[str, np.ndarray]:
    """The packed stream of a ``PackedEvents`` as numpy arrays."""
    return {
        "duration": packed.duration.detach().cpu().numpy(),
        "kind": packed.kind.cpu().numpy(),
        "flip": packed.flip.detach().cpu().numpy(),
        "phase": packed.phase.detach().cpu().numpy(),
        "action": packed.action.cpu().numpy(),
        "output_index": packed.output_index.cpu().numpy(),
        "saturation": packed.saturation.detach().cpu().numpy(),
        "rf_frequency": packed.rf_frequency_hz.detach().cpu().numpy(),
    }


def reference_module():
    """The reference's CPU extension built from source into ``oracle/_ref``, or None."""
    matches = sorted((HERE / "_ref").glob("_epg_cpu*.so"))
    if not matches:
        return None
    loader = importlib.machinery.ExtensionFileLoader("_epg_cpu", str(matches[0]))
    spec = importlib.util.spec_from_loader("_epg_cpu", loader)
    module = importlib.util.module_from_spec(spec)
    loader.exec_module(module)
    return module


_TISSUE_ORDER = (
    "t1", "t2", "m0", "b1", "b1_phase", "b0", "ie", "diffusion", "velocity",
    "bound_fraction", "bound_exchange", "t1_bound", "pool_b_fraction",
    "pool_b_exchange", "t1_pool_b", "t2_pool_b", "pool_b_shift",
)
_TISSUE_DEFAULT = {"m0": 1.0, "b1": 1.0, "ie": 1.0, "t1_bound": 1.0, "t1_pool_b": 1.0, "t2_pool_b": 1.0}


def _reference_call(
    events: dict[str, Any],
    tissue: dict[str, Any],
    *,
    states: int,
    outputs: int,
    features: int,
    real_axis: int,
    pool_kind: int,
    tangent: dict[str, Any] | None,
    threads: int,
    lineshape: tuple[Any, Any, float] | None = None,
    profile: tuple[Any, Any, Any, Any, float] | None = None,
) -> np.ndarray:
    module = reference_module()
    if module is None:
        raise RuntimeError("oracle/_ref is not built")
    atoms = max(np.size(v) for v in tissue.values())
    keep = []

    def f32(value, size):
        array = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float32), (size,)).copy())
        keep.append(array)
        return array.ctypes.data

    tissue_ptrs = [f32(tissue.get(name, _TISSUE_DEFAULT.get(name, 0.0)), atoms) for name in _TISSUE_ORDER]
    count = int(np.size(events["kind"]))

    def arr(value, dtype):
        array = np.ascontiguousarray(np.asarray(value, dtype=dtype))
        keep.append(array)
        return array.ctypes.data

    event_ptrs = [
        arr(events["duration"], np.float32),
        arr(events["kind"], np.int32),
        arr(events["flip"], np.float32),
        arr(events.get("phase", np.zeros(count)), np.float32),
        arr(events["action"], np.uint8),
        arr(events["output_index"], np.int32),
        arr(np.zeros(count), np.int32),
        arr(events.get("saturation", np.zeros(count)), np.float32),
        arr(events.get("rf_frequency", np.zeros(count)), np.float32),
    ]
    real = np.zeros((atoms, outputs), dtype=np.float32)
    imag = np.zeros((atoms, outputs), dtype=np.float32)
    tail = [0] * 13
    bins, step = 0, 1.0
    if lineshape is not None:
        # the absorption table sits in the last tail slot (Tail::ABSORPTION)
        knots = np.stack((np.asarray(lineshape[0]), np.asarray(lineshape[1])), -1)
        tail[12] = arr(knots, np.float32)
        bins, step = int(knots.shape[0]), float(lineshape[2])
    locations, profile_bins, profile_step = 1, 0, 1.0
    if profile is not None:
        # the table (positions, bins, 8) in Tail::TABLE and every event's
        # shape row (one shape: 0) in Tail::TABLE_INDEX
        a, b, da, db = (np.asarray(v, dtype=np.complex128) for v in profile[:4])
        table = np.stack((a.real, a.imag, b.real, b.imag, da.real, da.imag, db.real, db.imag), -1)
        tail[5] = arr(table, np.float32)
        tail[6] = arr(np.zeros(count), np.int32)
        locations, profile_bins = int(a.shape[0]), int(a.shape[1])
        profile_step = float(profile[4]) / (profile_bins - 1)
    dims = (atoms, 1, count, states, outputs, threads, real_axis, 0.0, 0.0, 1, locations, profile_bins,
            profile_step, bins, step, pool_kind, features, 0, 0, 0)
    if tangent is None:
        module.simulate((*tissue_ptrs, *event_ptrs, real.ctypes.data, imag.ctypes.data, *tail), *dims)
    else:
        tangent_ptrs = [f32(tangent.get(name, 0.0), atoms) for name in _TISSUE_ORDER]
        # event tangents in the reference's forward order: duration, flip, phase
        tangent_ptrs += [f32(tangent.get(name, 0.0), count) for name in ("duration", "flip", "phase")]
        module.simulate_jvp(
            (*tissue_ptrs, *event_ptrs, *tangent_ptrs, real.ctypes.data, imag.ctypes.data, *tail), *dims
        )
    return real.astype(np.float64) + 1j * imag.astype(np.float64)


#: Feature bits of the reference's packed ABI (src/torchsim/_epg_cpu.cpp:90-104).
FEATURE_ALL = 63


def reference_real(
    events: dict[str, Any],
    tissue: dict[str, Any],
    *,
    states: int,
    outputs: int,
    features: int,
    tangent: dict[str, Any] | None = None,
    threads: int = 1,
    real_axis: int = 1,
) -> np.ndarray:
    """Run the reference's own FP32 CPU kernel (built from source) on packed buffers.

    Returns the imaginary output plane (atoms, outputs): the recorded
    ``m0 * a0`` for the forward, or its tangent along ``tangent`` (per-atom
    rows keyed like ``tissue``) through ``simulate_jvp``.
    """
    out = _reference_call(events, tissue, states=states, outputs=outputs, features=features,
                          real_axis=real_axis, pool_kind=0, tangent=tangent, threads=threads)
    return out.imag.copy()


def reference_complex(
    events: dict[str, Any],
    tissue: dict[str, Any],
    *,
    states: int,
    outputs: int,
    features: int = FEATURE_ALL,
    exchanging: bool = False,
    tangent: dict[str, Any] | None = None,
    threads: int = 1,
    lineshape: tuple[Any, Any, float] | None = None,
    profile: tuple[Any, Any, Any, Any, float] | None = None,
) -> np.ndarray:
    """The reference's own FP32 complex kernel: recorded samples (atoms, outputs), complex.

    Runs ``simulate`` (or ``simulate_jvp`` along ``tangent``) with
    ``real_axis = -1``, i.e. the general complex state machine
    (src/torchsim/_epg_cpu.cpp:4216-4566, dispatch :10789-10900), with pool
    kind 2 -- the chemically exchanging second pool
    (:885-932, :1714-1764) -- when ``exchanging``. ``tissue`` is keyed by the
    reference's 17 ABI names (``t1, t2, m0, b1, b1_phase, b0, ie, diffusion,
    velocity, bound_fraction, bound_exchange, t1_bound, pool_b_fraction,
    pool_b_exchange, t1_pool_b, t2_pool_b, pool_b_shift``); ``diffusion`` is
    the rate already scaled to 1/s per unit order squared. ``tangent`` takes
    the same names plus per-event ``duration``, ``flip`` and ``phase``.
    With ``lineshape`` = (values, slopes, step_hz) the tissue's
    ``bound_fraction`` / ``bound_exchange`` / ``t1_bound`` run the semisolid
    pool (pool kind 1), the events' ``saturation`` and ``rf_frequency``
    driving its saturation. With ``profile`` = (a, b, slope_a, slope_b,
    theta_max) ((positions, bins) each) the pulses are profiled (RfMode::
    PROFILED): ``tissue`` then holds one row per (voxel, position),
    voxel-major.
    """
    kind = 1 if lineshape is not None else (2 if exchanging else 0)
    return _reference_call(events, tissue, states=states, outputs=outputs, features=features,
                           real_axis=-1, pool_kind=kind, tangent=tangent, threads=threads, lineshape=lineshape,
                           profile=profile)


_EVENT_VJP_ORDER = ("flip", "phase", "duration")        # simulate_vjp's gradient slots after the tissue
_EVENT_JVP_ORDER = ("duration", "flip", "phase")        # simulate_jvp / simulate_vjp_jvp's


def reference_vjp(
    events: dict[str, Any],
    tissue: dict[str, Any],
    cotangent: np.ndarray,
    *,
    states: int,
    outputs: int,
    features: int = FEATURE_ALL,
    real_axis: int = 1,
    exchanging: bool = False,
    tangent: dict[str, Any] | None = None,
    threads: int = 8,
    lineshape: tuple[Any, Any, float] | None = None,
) -> dict[str, np.ndarray]:
    """Reverse mode of the reference's own FP32 kernel (built from source).

    ``cotangent`` is complex (atoms, outputs): its real and imaginary parts
    seed the two output planes. Without ``tangent`` this is
    ``simulate_vjp`` (src/torchsim/_epg_cpu.cpp:11150-11416): the gradient
    of <w, y(x)>. With ``tangent`` (per-atom tissue rows and per-event
    ``duration``/``flip``/``phase`` rows) it is ``simulate_vjp_jvp``
    (:11589-11755), and the result is the gradient of <w, J(x) v> -- the
    forward-over-reverse pass ``_NativeEpgJvp.backward`` takes
    (sequence/_accelerators.py:1787-1850). Returns the 17 tissue gradients
    (atoms,) by name and the event gradients ``flip``, ``phase``,
    ``duration`` (events,), summed over atoms. With ``lineshape`` =
    (values, slopes, step_hz) the tissue's bound_fraction / bound_exchange /
    t1_bound form a semisolid pool (pool kind 1, the table in
    Tail::ABSORPTION) saturated by the events' ``saturation`` and
    ``rf_frequency``.
    """
    module = reference_module()
    if module is None:
        raise RuntimeError("oracle/_ref is not built")
    atoms = max(np.size(v) for v in tissue.values())
    count = int(np.size(events["kind"]))
    keep = []

    def f32(value, size):
        array = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=np.float32), (size,)).copy())
        keep.append(array)
        return array

    def arr(value, dtype):
        array = np.ascontiguousarray(np.asarray(value, dtype=dtype))
        keep.append(array)
        return array.ctypes.data

    tissue_ptrs = [f32(tissue.get(n, _TISSUE_DEFAULT.get(n, 0.0)), atoms).ctypes.data for n in _TISSUE_ORDER]
    event_ptrs = [
        arr(events["duration"], np.float32), arr(events["kind"], np.int32), arr(events["flip"], np.float32),
        arr(events.get("phase", np.zeros(count)), np.float32), arr(events["action"], np.uint8),
        arr(events["output_index"], np.int32), arr(np.zeros(count), np.int32),
        arr(events.get("saturation", np.zeros(count)), np.float32),
        arr(events.get("rf_frequency", np.zeros(count)), np.float32),
    ]
    w = np.asarray(cotangent).reshape(atoms, outputs)
    seed = [f32(w.real.reshape(-1), atoms * outputs).ctypes.data, f32(w.imag.reshape(-1), atoms * outputs).ctypes.data]
    tail = [0] * 13
    bins, step = 0, 1.0
    if lineshape is not None:
        knots = np.stack((np.asarray(lineshape[0]), np.asarray(lineshape[1])), -1)
        tail[12] = arr(knots, np.float32)
        bins, step = int(knots.shape[0]), float(lineshape[2])
    kind = 1 if lineshape is not None else (2 if exchanging else 0)
    dims = (atoms, 1, count, states, outputs, threads, real_axis, 0.0, 0.0, 1, 1, 0, 1.0, bins, step,
            kind, features, 0, 0, 0)

    def gradient_block(order):
        tissue_grads = [f32(0.0, atoms) for _ in _TISSUE_ORDER]
        event_grads = {name: f32(0.0, count) for name in order}
        pointers = [g.ctypes.data for g in tissue_grads] + [event_grads[n].ctypes.data for n in order]
        out = {name: g for name, g in zip(_TISSUE_ORDER, tissue_grads)}
        out.update(event_grads)
        return pointers, out

    if tangent is None:
        pointers, grads = gradient_block(_EVENT_VJP_ORDER)
        module.simulate_vjp((*tissue_ptrs, *event_ptrs, *seed, *pointers, *tail), *dims)
    else:
        direction = [f32(tangent.get(n, 0.0), atoms).ctypes.data for n in _TISSUE_ORDER]
        direction += [f32(tangent.get(n, 0.0), count).ctypes.data for n in _EVENT_JVP_ORDER]
        value_pointers, _ = gradient_block(_EVENT_JVP_ORDER)
        pointers, grads = gradient_block(_EVENT_JVP_ORDER)
        module.simulate_vjp_jvp((*tissue_ptrs, *event_ptrs, *direction, *seed, *value_pointers, *pointers, *tail),
                                *dims)
    return {name: value.astype(np.float64) for name, value in grads.items()}


def reference_crlb_gradient(
    events: dict[str, Any],
    t1: Any,
    t2: Any,
    *,
    states: int,
    outputs: int,
    factor: complex = 1j,
    threads: int = 8,
) -> tuple[float, np.ndarray, np.ndarray]:
    """The sequence-design loss sum_atoms tr(F^-1) of (T1, T2, M0) at sigma 1
    and its gradient with respect to every event's flip and duration, the way
    the reference's autograd reaches it: Jacobian columns from its FP32
    kernel (``simulate_jvp``; the M0 column is the signal at M0 = 1), the loss
    and its cotangents on the columns in FP64, then one forward-over-reverse
    pass per tissue column (``simulate_vjp_jvp``) and one first-order pass
    for the M0 column (``simulate_vjp``), summed -- what
    ``crlb(J).sum().backward()`` runs through ``_NativeEpgJvp.backward`` and
    ``_NativeEpg.backward`` (optim/_design.py:196-253,
    sequence/_accelerators.py:1544-1586, 1787-1850).

    ``events`` is a real-subspace stream (the kernels' (0, m0 a0) planes);
    ``factor`` is the simulator's constant (MRF: 1j), so a public column is
    factor * i * (its imaginary plane). Returns (loss, d/dflip, d/dduration)
    with the event gradients (events,).
    """
    t1 = np.asarray(t1, dtype=np.float32)
    t2 = np.asarray(t2, dtype=np.float32)
    tissue = {"t1": t1, "t2": t2, "m0": np.ones_like(t1)}
    kw = dict(states=states, outputs=outputs, features=FEATURE_ALL, threads=threads)
    signal = reference_real(events, tissue, **kw)
    columns = [reference_real(events, tissue, tangent={name: 1.0}, **kw) for name in ("t1", "t2")] + [signal]
    scale = complex(factor) * 1j                     # public = scale * plane
    jac = np.stack(columns, 1).astype(np.float64) * scale
    rows = np.concatenate((jac.real, jac.imag), -1)   # (atoms, 3, 2 outputs)
    fisher = rows @ np.transpose(rows, (0, 2, 1))
    inverse = np.linalg.inv(fisher)
    loss = float(np.trace(inverse, axis1=1, axis2=2).sum())
    d_rows = -2.0 * (inverse @ inverse) @ rows        # dL/d rows
    d_public = d_rows[..., :outputs] + 1j * d_rows[..., outputs:]
    # <w_pub, scale * plane> over the real pair (re, im) -> cotangent on the plane
    d_plane = (np.conj(scale) * d_public).real        # real, (atoms, 3, outputs)
    flip = np.zeros(np.size(events["kind"]))
    duration = np.zeros_like(flip)
    for k, name in enumerate(("t1", "t2")):
        g = reference_vjp(events, tissue, 1j * d_plane[:, k], tangent={name: 1.0}, **kw)
        flip += g["flip"]
        duration += g["duration"]
    g = reference_vjp(events, tissue, 1j * d_plane[:, 2], **kw)
    flip += g["flip"]
    duration += g["duration"]
    return loss, flip, duration
