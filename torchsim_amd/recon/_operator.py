"""A signal model over bounded unknowns, as fused normal equations per voxel.

The role of the reference's ModelOperator (recon/_operator.py:342-483) for a
model with nothing encoding the voxels together: hold the simulator and the
unknowns, map free variables to the properties, and give the loop what it
needs per voxel. Where the reference hands back the Jacobian blocks, this
hands back cost, J^T r and J^T J from one gfx950 launch.
"""

from __future__ import annotations

import ctypes
from collections.abc import Mapping, Sequence
from typing import Any

import torch

from .. import _native
from ..sequence import _engine
from ._bounds import to_free, to_natural, widen

_COLUMN_INDEX = {"t1_ms": 0, "t2_ms": 1, "m0": 2, "b1": 3}


class FusedModel:
    """``simulator`` evaluated at per-voxel unknowns ``names`` (public property
    names, e.g. ``("T1", "T2", "M0")``), bounded by ``bounds``."""

    def __init__(self, simulator: Any, names: Sequence[str], *, bounds: Mapping[str, Any] | None = None,
                 known: Mapping[str, Any] | None = None, subspace: Any = None) -> None:
        self.simulator = simulator
        # with a subspace, data and model are (voxels, rank) coefficients
        # (reference recon/_operator.py:451-473 projects both)
        self.subspace = subspace
        self.names = tuple(names)
        self.bounds = dict(bounds or {})
        self.known = dict(known or {})
        held, sequence = simulator._split({**simulator.bound, **self.known})
        self.sequence = sequence
        self.held = {name: value for name, value in held.items() if name not in self.names}
        self.fields = tuple(simulator._tissue_fields({name: 1.0 for name in self.names}))
        for field in self.fields:
            if field not in _COLUMN_INDEX:
                raise NotImplementedError(f"the fused fit carries no derivative along {field!r}")
        self.packed = simulator.packed(**sequence)
        self.states = int(sequence.get("states", simulator.states) or 10)

    def held_on(self, device: torch.device) -> dict[str, Any]:
        """The fixed properties as the kernels take them: per-voxel tensors
        moved to ``device`` once, numbers and host scalars left as they are
        (they become device fills, with no wait on the stream)."""
        cache = self.__dict__.setdefault("_held_device", {})
        key = str(device)
        if key not in cache:
            cache[key] = {name: value.to(device) if isinstance(value, torch.Tensor) and value.numel() > 1 else value
                          for name, value in self.held.items()}
        return cache[key]

    @property
    def columns(self) -> int:
        return len(self.names)

    def natural(self, x: torch.Tensor) -> torch.Tensor:
        return to_natural(x, self.bounds, self.names)

    def free(self, natural: torch.Tensor) -> torch.Tensor:
        return to_free(natural, self.bounds, self.names)

    def scale(self, x: torch.Tensor) -> torch.Tensor:
        return widen(self.bounds, self.names, x).to(torch.float64)

    def normal(self, x: torch.Tensor, data: torch.Tensor, live: torch.Tensor | None) -> torch.Tensor:
        """(voxels, stride) FP64: cost, J^T r (natural columns), J^T J, for the
        voxels ``live`` indexes into per-voxel fixed properties (None: all);
        ``data`` holds those voxels' measured rows, in the same order."""
        values = self.natural(x)
        return self.normal_at([values[:, i].to(torch.float32) for i in range(self.columns)], data, live)

    def normal_at(self, columns: Sequence[torch.Tensor], data: torch.Tensor,
                  live: torch.Tensor | None) -> torch.Tensor:
        """As :meth:`normal`, at the unknowns' natural values given as one
        float32 tensor per unknown."""
        device = data.device
        fields: dict[str, Any] = {}
        for name, value in self.held_on(device).items():
            if isinstance(value, torch.Tensor) and value.numel() > 1:
                value = value.reshape(-1) if live is None else value.reshape(-1)[live]
            fields.update(self.simulator._tissue_fields({name: value}))
        for index, field in enumerate(self.fields):
            fields[field] = columns[index]
        if self.subspace is not None or self.__dict__.get("_unfused"):
            return self._normal_from_jacobian(columns, data, live)
        tissue = _engine.prepare_tissue(fields, device, self.fields)
        real, flip, samples = _engine.real_subspace(self.packed, tissue, device)
        if not real or tissue.features & _native.FEATURE_DIFFUSING or self.simulator.signal_factor is None:
            # the verdict depends on the protocol and the fixed properties only
            self._unfused = True
            return self._normal_from_jacobian(columns, data, live)
        lib = _native.load()
        events = list(self.packed.on(device))
        if flip is not None:
            events[2] = flip
        stride = lib.tsim_normal_stride()
        normal = torch.empty((tissue.atoms, stride), dtype=torch.float64, device=device)
        if tissue.atoms == 0:
            return normal
        measured = data.to(device=device, dtype=torch.complex64).contiguous()
        if measured.shape[0] != tissue.atoms:
            raise ValueError(f"data holds {measured.shape[0]} rows for {tissue.atoms} voxels")
        cols = (ctypes.c_int32 * self.columns)(*[_COLUMN_INDEX[f] for f in self.fields])
        sign = samples
        factor = complex(self.simulator.signal_factor)

        def launch(held: _engine.PreparedTissue, rows: torch.Tensor, out: torch.Tensor, tile: int) -> None:
            with torch.cuda.device(device):
                status = lib.tsim_gauss_newton_normal(
                    _native.pointers([*held.buffers, *events]), held.atoms, self.packed.event_count, self.states,
                    self.packed.output_count, held.features, held.stride, cols, self.columns, factor.real,
                    factor.imag, None if sign is None else sign.data_ptr(), rows.data_ptr(), out.data_ptr(), tile,
                    _native.stream_of(device),
                )
            _native.check(status)

        launch(tissue, measured, normal, self._normal_tile(tissue, measured, stride, launch))
        return normal

    def _normal_tile(self, tissue: Any, measured: torch.Tensor, stride: int, launch: Any) -> int:
        """The state tile of the fused normal equations, measured once per
        configuration and size class (every tile sums the same arithmetic in
        the same order, so the choice changes only the speed). A tile with
        more lanes per voxel fills more of the chip when voxels are few and
        carries more idle lanes through the per-voxel FP64 sums when they are
        many, so launches of 65,536 voxels or more are probed on up to
        262,144 of theirs and kept apart from smaller ones (probed on up to
        16,384)."""
        from ..sequence import _tuning

        large = tissue.atoms >= 4 * _tuning.PROBE_ATOMS
        probe_tissue = _engine.narrow_tissue(tissue, (16 if large else 1) * _tuning.PROBE_ATOMS)
        if probe_tissue.atoms == 0:
            return -1
        rows = measured[:probe_tissue.atoms]
        out = torch.empty((probe_tissue.atoms, stride), dtype=torch.float64, device=measured.device)
        return _tuning.tile_for(measured.device, self.states, tuple(self.fields), tissue.features,
                                lambda tile: launch(probe_tissue, rows, out, tile),
                                tag="normal equations, " + ("large" if large else "small"))

    def _normal_from_jacobian(self, columns: Sequence[torch.Tensor], data: torch.Tensor,
                              live: torch.Tensor | None) -> torch.Tensor:
        """Any sequence: the simulator's fused signal + Jacobian (complex
        kernel, post-processing included), reduced to the normal equations in
        FP64 by tsim_normal_from_jacobian -- the Jacobian passes through memory
        once on this path. With a subspace, signal and Jacobian arrive as
        coefficients, projected inside the kernel where the model allows."""
        device = data.device
        held = {}
        for name, value in self.held_on(device).items():
            if isinstance(value, torch.Tensor) and value.numel() > 1:
                value = value.reshape(-1) if live is None else value.reshape(-1)[live]
            held[name] = value
        for index, name in enumerate(self.names):
            held[name] = columns[index]
        if self.subspace is not None:
            signal, jacobian = self.simulator.linearize_projected(held, self.names, self.subspace, **self.sequence)
        else:
            signal, jacobian = self.simulator.linearize(held, self.names, **self.sequence)
        lib = _native.load()
        stride = lib.tsim_normal_stride()
        voxels, outputs = signal.shape
        normal = torch.zeros((voxels, stride), dtype=torch.float64, device=device)
        if voxels == 0:
            return normal
        signal = signal.to(torch.complex64).contiguous()
        jacobian = jacobian.to(torch.complex64).contiguous()
        measured = data.to(device=device, dtype=torch.complex64).contiguous()
        with torch.cuda.device(device):
            status = lib.tsim_normal_from_jacobian(
                signal.data_ptr(), jacobian.data_ptr(), measured.data_ptr(), voxels, outputs, self.columns,
                normal.data_ptr(), _native.stream_of(device),
            )
        _native.check(status)
        return normal
