"""Host side of the uneven-batch launcher: many requests, one launch."""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from .. import _native
from ..sequence._engine import PreparedTissue
from ..sequence._events import PackedEvents

_COLUMN_INDEX = {"t1_ms": 0, "t2_ms": 1, "m0": 2, "b1": 3, "inversion_efficiency": 6}

#: Jacobian columns the complex kernel's batch carries.
COMPLEX_COLUMNS = ("t1_ms", "t2_ms", "m0", "b1")


@dataclass
class BatchItem:
    """One request's share of a batched launch (all device tensors)."""

    packed: PackedEvents
    tissue: PreparedTissue
    states: int
    columns: tuple[str, ...] = ()
    factor: complex = 1.0
    flip: torch.Tensor | None = None
    samples: torch.Tensor | None = None
    signal: torch.Tensor | None = None      # mode 0 output
    jacobian: torch.Tensor | None = None    # mode 0 output
    data: torch.Tensor | None = None        # mode 1 input
    normal: torch.Tensor | None = None      # mode 1 output
    kernel: str = "real"                    # "real" subspace or "complex"

    def variant(self) -> tuple:
        """What decides the kernel a request runs: tangents, diffusion, transmit
        (for the complex kernel the tangents alone: its features, pool count
        and state count are per request)."""
        if self.kernel == "complex":
            mask = 0
            for name in self.columns:
                mask |= {"t1_ms": 1, "t2_ms": 2, "b1": 4}.get(name, 0)
            return "complex", mask
        mask = 0
        for name in self.columns:
            mask |= {"t1_ms": 1, "t2_ms": 2, "b1": 4, "inversion_efficiency": 8}.get(name, 0)
        if mask & 8:
            mask = 15
        return mask, bool(self.tissue.features & _native.FEATURE_DIFFUSING), bool(
            self.tissue.features & _native.FEATURE_TRANSMIT)


def launch_batch(items: list[BatchItem], *, mode: int, stream: torch.cuda.Stream, tile: int = -1,
                 grid_blocks: int = 0) -> None:
    """Launch every item in one persistent kernel on ``stream`` (asynchronous).

    Items must share one kernel variant (see :meth:`BatchItem.variant`). The
    caller keeps every tensor alive until the stream has passed the launch.
    Complex-kernel items (mode 0 only) take ``tsim_complex_batch``, whose
    grid walks the same work-block table, one request's pool model per block.

    The launch can be captured into a HIP graph (``torch.cuda.graph``) and
    replayed: run it eagerly, synchronize and run it once more before the
    capture (event buffers and the pinned table staging then exist, and
    nothing is allocated on the host while the stream captures); the graph
    keeps its copy of the request and work tables, re-reads the items'
    tissue and writes their outputs at every replay.
    """
    if not items:
        return
    lib = _native.load()
    complex_kernel = items[0].kernel == "complex"
    if any((item.kernel == "complex") != complex_kernel for item in items):
        raise ValueError("a batch runs one kernel: real-subspace and complex requests go in separate batches")
    if complex_kernel and mode != 0:
        raise ValueError("the complex kernel's batch writes signals and Jacobians (mode 0)")
    table = (_native.BatchRequest * len(items))()
    keep = []
    device = stream.device
    for entry, item in zip(table, items):
        with torch.cuda.stream(stream):
            # (the launch stream waits for the event buffers' upload, if any)
            events = list(item.packed.on(device))
        if item.flip is not None:
            events[2] = item.flip
        buffers = [*item.tissue.buffers, *events]
        for k, tensor in enumerate(buffers):
            entry.packed[k] = None if tensor is None else tensor.data_ptr()
        entry.atom_count = item.tissue.atoms
        entry.event_count = item.packed.event_count
        entry.state_count = item.states
        entry.output_count = item.packed.output_count
        entry.tissue_stride = item.tissue.stride
        entry.features = item.tissue.features
        entry.column_count = len(item.columns)
        for k, name in enumerate(item.columns):
            if complex_kernel and name not in COMPLEX_COLUMNS:
                raise ValueError(f"the complex kernel carries no Jacobian column along {name!r}")
            entry.columns[k] = _COLUMN_INDEX[name]
        factor = complex(item.factor)
        entry.factor_re = factor.real
        entry.factor_im = factor.imag
        entry.sample_sign = None if item.samples is None else item.samples.data_ptr()
        entry.signal = None if item.signal is None else item.signal.data_ptr()
        entry.jacobian = (None if item.jacobian is None or (complex_kernel and not item.columns)
                          else item.jacobian.data_ptr())
        entry.data = None if item.data is None else item.data.data_ptr()
        entry.normal = None if item.normal is None else item.normal.data_ptr()
        keep.append(buffers)
        # the launch stream reads and writes memory other streams allocated
        for tensor in [*buffers, item.samples, item.signal, item.jacobian, item.data, item.normal]:
            if tensor is not None and tensor.is_cuda:
                tensor.record_stream(stream)
    with torch.cuda.device(device):
        if complex_kernel:
            status = lib.tsim_complex_batch(table, len(items), grid_blocks, stream.cuda_stream)
        else:
            status = lib.tsim_simulate_batch(table, len(items), mode, tile, grid_blocks, stream.cuda_stream)
    _native.check(status)
