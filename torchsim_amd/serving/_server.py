"""A continuous-batching request loop for simulate and fit requests.

No reference counterpart: torchsim is a library with no request loop. This
serves many callers on one MI355X:

* **Admission by HBM budget.** Each request states (estimates) the device
  bytes it will hold -- inputs, outputs, fit state -- and is admitted when
  that fits in what the budget has left; a request larger than the whole
  budget is refused at submission.
* **Continuous batching.** The scheduler works in ticks. At every tick each
  active request offers its next kernel job (a simulate request one, a fit
  one per Gauss-Newton iteration); jobs that run the same kernel variant are
  packed into ONE uneven-batch launch (``tsim_simulate_batch``) whatever
  their sizes, sequences and state counts. Requests that arrive while others
  are on the GPU join at the next tick; nothing waits for a batch to drain.
* **Overlap.** Each request keeps its own HIP stream for its own work; a
  batched launch runs on a stream of its own that waits on its members'
  events and hands back through one event, so independent groups and the
  requests' elementwise updates overlap on the device.
* **Residency.** Fit state (free variables, damping, normal equations, the
  measured data) stays in HBM across iterations; a simulate request may ask
  to keep its result resident under a name (charged to the budget until
  released).
* **Same alone and in the mix.** Every atom of a batched launch runs the
  arithmetic its request alone would (the kernel is per-atom), and the
  Gauss-Newton update is per voxel, so a request's result does not depend on
  what it was batched with.

Requests whose kernel is not batchable (complex sequences, diffusion in a
fit) run on their own stream through the ordinary simulator path.
"""

from __future__ import annotations

import itertools
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Any

import torch

from .. import _native
from ..recon._gauss_newton import solve_damped
from ..sequence import _engine
from ._batch import BatchItem, launch_batch

_TINY = torch.finfo(torch.float32).tiny


@dataclass
class SimulateRequest:
    """Simulate ``simulator`` (protocol bound on it) at ``properties``; with
    ``diff``, also its Jacobian. ``resident`` keeps the result on the device
    under that name."""

    simulator: Any
    properties: dict[str, Any]
    diff: tuple[str, ...] = ()
    resident: str | None = None


@dataclass
class FitRequest:
    """Fit ``unknown`` (public names) of ``simulator`` to ``measured``
    (voxels, samples) by Levenberg-Marquardt, ``iterations`` steps from
    ``initial`` inside ``bounds``; ``known`` fixes other properties."""

    simulator: Any
    measured: torch.Tensor
    unknown: tuple[str, ...] = ("T1", "T2", "M0")
    bounds: dict[str, tuple[float, float]] = field(default_factory=dict)
    initial: dict[str, float] = field(default_factory=dict)
    known: dict[str, Any] = field(default_factory=dict)
    iterations: int = 10
    tau: float = 1e-2


@dataclass
class _Ticket:
    request: Any
    future: Future
    arrival: float
    bytes: int
    program: Any = None
    stream: torch.cuda.Stream | None = None
    last: torch.cuda.Event | None = None
    admitted: float = 0.0
    done_event: torch.cuda.Event | None = None


class _SimulateProgram:
    """One batched job (real subspace), or the ordinary path on its stream."""

    def __init__(self, request: SimulateRequest, device: torch.device) -> None:
        sim = request.simulator
        held, sequence = sim._split({**sim.bound, **request.properties})
        self.held, self.sequence, self.sim, self.names = held, sequence, sim, tuple(request.diff)
        self.batch_item: BatchItem | None = None
        self.result: Any = None
        self.launched = False
        fields = sim._tissue_fields(held)
        differentiated = tuple(sim._tissue_fields({n: held[n] for n in self.names}))
        packed = sim.packed(**sequence)
        batchable = sim.signal_factor is not None and "diffusion_um2_per_ms" not in fields
        if batchable:
            tissue = _engine.prepare_tissue(fields, device, differentiated)
            real, flip, samples = _engine.real_subspace(packed, tissue, device)
            if real:
                states = int(sequence.get("states", sim.states) or 10)
                atoms, n = tissue.atoms, packed.output_count
                signal = torch.empty((atoms, n), dtype=torch.complex64, device=device)
                jacobian = torch.empty((atoms, len(differentiated), n), dtype=torch.complex64, device=device)
                self.batch_item = BatchItem(packed, tissue, states, differentiated, sim.signal_factor,
                                            flip=flip, samples=samples, signal=signal, jacobian=jacobian)
                self.batch = sim._batch(held)

    def job(self) -> tuple[int, BatchItem] | None:
        if self.batch_item is None or self.launched:
            return None
        return 0, self.batch_item

    def run_unbatched(self) -> None:
        if self.names:
            self.result = self.sim.jacobian(self.names, **{**self.held, **self.sequence})
        else:
            self.result = self.sim.simulate(**{**self.held, **self.sequence})
        self.launched = True

    def after_job(self) -> None:
        self.launched = True
        item = self.batch_item
        n = item.signal.shape[-1]
        signal = item.signal.reshape(*self.batch, n)
        if self.names:
            self.result = (signal, item.jacobian.reshape(*self.batch, len(self.names), n))
        else:
            self.result = signal

    @property
    def finished(self) -> bool:
        return self.launched


class _FitProgram:
    """Levenberg-Marquardt over a fixed number of iterations, all on device.

    Same per-voxel algorithm as GaussNewton(TrustRegion()) -- Nielsen damping,
    accept on positive gain -- with a fixed iteration count and no host reads
    in between, so every iteration is one batched normal-equation job plus a
    few asynchronous elementwise updates.
    """

    def __init__(self, request: FitRequest, device: torch.device, graphs: bool = False) -> None:
        from ..recon import FusedModel

        self.request = request
        self.graphs = graphs
        self._graph: torch.cuda.CUDAGraph | None = None
        self._graph_failed = False
        self.op = FusedModel(request.simulator, request.unknown, bounds=request.bounds, known=request.known)
        self.data = request.measured.to(device=device, dtype=torch.complex64).reshape(
            -1, request.measured.shape[-1]).contiguous()
        voxels = self.data.shape[0]
        self.live = None  # every voxel iterates (no retirement in this loop)
        # the start as device fills (no host-to-device copy, no wait)
        start = torch.stack([torch.full((), float(request.initial.get(n, 1.0)), dtype=torch.float64, device=device)
                             for n in request.unknown])
        self.x = self.op.free(start[None, :].expand(voxels, -1)).contiguous()
        self.iteration = -1  # -1: the start point still needs its normal equations
        self.normal = None
        self.candidate = None
        self.cost = None
        self.damping = None
        self.rising = None
        self.step_predicted = None
        self.fused = self._fusable()
        self.shape = request.measured.shape[:-1]

    def _fusable(self) -> bool:
        try:
            fields = {}
            for name, value in self.op.held_on(self.data.device).items():
                fields.update(self.op.simulator._tissue_fields({name: value}))
            if "diffusion_um2_per_ms" in fields or self.op.simulator.signal_factor is None:
                return False
            probe = {f: 1.0 for f in self.op.fields}
            tissue = _engine.prepare_tissue({**fields, **probe}, self.data.device, self.op.fields)
            return _engine.real_subspace(self.op.packed, tissue)[0]
        except (NotImplementedError, ValueError):
            return False

    def _item(self, x: torch.Tensor) -> BatchItem:
        device = self.data.device
        values = self.op.natural(x)
        fields: dict[str, Any] = {}
        for name, value in self.op.held_on(device).items():
            fields.update(self.op.simulator._tissue_fields({name: value}))
        for index, f in enumerate(self.op.fields):
            fields[f] = values[:, index].to(torch.float32)
        tissue = _engine.prepare_tissue(fields, device, self.op.fields)
        real, flip, samples = _engine.real_subspace(self.op.packed, tissue, device)
        stride = _native.load().tsim_normal_stride()
        normal = torch.empty((tissue.atoms, stride), dtype=torch.float64, device=device)
        return BatchItem(self.op.packed, tissue, self.op.states, self.op.fields, self.op.simulator.signal_factor,
                         flip=flip, samples=samples, data=self.data, normal=normal)

    def job(self) -> tuple[int, BatchItem] | None:
        if self.finished or not self.fused:
            return None
        if self.iteration < 0:
            self._pending = self._item(self.x)
        else:
            self._prepare_step()
            self._pending = self._item(self.candidate)
        return 1, self._pending

    def _prepare_step(self) -> None:
        columns = self.op.columns
        step, predicted, _, _ = solve_damped(self.normal, self.op.scale(self.x), self.damping, columns)
        singular = ~torch.isfinite(step).all(-1)
        self.singular = singular
        self.step_predicted = predicted
        self.candidate = torch.where(singular[:, None], self.x, self.x + torch.nan_to_num(step))

    def after_job(self, normal: torch.Tensor | None = None) -> None:
        normal = self._pending.normal if normal is None else normal
        if self.iteration < 0:
            self.normal = normal.clone()
            self.cost = normal[:, 0].clone()
            _, _, _, curvature = solve_damped(normal, self.op.scale(self.x), None, self.op.columns, step=False)
            self.damping = self.request.tau * curvature
            self.rising = torch.full_like(self.damping, 2.0)
        else:
            trial_cost = normal[:, 0]
            predicted = self.step_predicted
            gain = torch.where(predicted > 0, (self.cost - trial_cost) / predicted.clamp_min(_TINY),
                               torch.zeros_like(predicted))
            better = (gain > 0) & ~self.singular
            lowered = self.damping * (1.0 - (2.0 * gain - 1.0).pow(3)).clamp_min(1.0 / 3.0)
            rising = self.rising
            self.rising = torch.where(better, torch.full_like(rising, 2.0), rising * 2.0)
            self.damping = torch.where(better, lowered, self.damping * rising)
            self.x = torch.where(better[:, None], self.candidate, self.x)
            self.normal = torch.where(better[:, None], normal, self.normal)
            self.cost = torch.where(better, trial_cost, self.cost)
        self.iteration += 1

    def run_unbatched(self) -> None:
        """Any sequence: one iteration of the same algorithm, with the normal
        equations formed from the simulator's fused signal + Jacobian (the
        scheduler calls this once per tick, so other requests interleave).

        The start point and the first iteration run eagerly; the iteration is
        then captured once into a HIP graph on the request's stream and
        replayed for the rest, so an iteration costs the host one replay
        instead of ~100 operator dispatches. Replays run exactly the kernels
        the eager iteration runs, on the same state tensors, updated in place.
        """
        if self.finished:
            return
        if self.iteration < 0:
            normal = self.op.normal(self.x, self.data, self.live)
            self._pending = None
            self.after_job(normal)
            return
        if self._graph is None and self.graphs and not self._graph_failed and self.iteration >= 1 \
                and self._uploads_settled():
            self._capture()
        if self._graph is not None:
            self._graph.replay()
        else:
            self._iterate_in_place()
        self.iteration += 1

    def release(self) -> None:
        """Drop the captured iteration and its memory. Only once the GPU has
        passed the request's last work: the graph's buffers (its private
        pool, its allocation nodes) must outlive every replay in flight."""
        self._graph = None

    def _iterate_in_place(self) -> None:
        """One Levenberg-Marquardt iteration on the state tensors (x, normal
        equations, cost, damping, its growth), written back in place."""
        step, predicted, _, _ = solve_damped(self.normal, self.op.scale(self.x), self.damping, self.op.columns)
        singular = ~torch.isfinite(step).all(-1)
        candidate = torch.where(singular[:, None], self.x, self.x + torch.nan_to_num(step))
        normal = self.op.normal(candidate, self.data, self.live)
        trial_cost = normal[:, 0]
        gain = torch.where(predicted > 0, (self.cost - trial_cost) / predicted.clamp_min(_TINY),
                           torch.zeros_like(predicted))
        better = (gain > 0) & ~singular
        lowered = self.damping * (1.0 - (2.0 * gain - 1.0).pow(3)).clamp_min(1.0 / 3.0)
        rising = torch.where(better, torch.full_like(self.rising, 2.0), self.rising * 2.0)
        damping = torch.where(better, lowered, self.damping * self.rising)
        x = torch.where(better[:, None], candidate, self.x)
        normal = torch.where(better[:, None], normal, self.normal)
        cost = torch.where(better, trial_cost, self.cost)
        self.x.copy_(x)
        self.normal.copy_(normal)
        self.cost.copy_(cost)
        self.damping.copy_(damping)
        self.rising.copy_(rising)

    def _uploads_settled(self) -> bool:
        """Whether the protocol's device buffers have landed (a graph cannot
        wait for an upload made outside it)."""
        uploads = self.op.packed.__dict__.get("_resident", {})
        for upload in uploads.values():
            if upload.ready is not None:
                if not upload.ready.query():
                    return False
                upload.ready = None
        return True

    def _capture(self) -> None:
        graph = torch.cuda.CUDAGraph()
        try:
            # thread-local: other threads (callers building requests) keep
            # allocating and synchronizing while this stream captures
            graph.capture_begin(capture_error_mode="thread_local")
            try:
                self._iterate_in_place()
            finally:
                graph.capture_end()
        except Exception:  # anything not capturable: stay eager for this request
            self._graph_failed = True
            return
        self._graph = graph

    @property
    def finished(self) -> bool:
        return self.iteration >= self.request.iterations

    @property
    def result(self) -> dict[str, torch.Tensor]:
        values = self.op.natural(self.x)
        return {n: values[:, i].reshape(self.shape).to(torch.float32) for i, n in enumerate(self.request.unknown)}


def _request_bytes(request: Any) -> int:
    """Device bytes a request holds while active (estimate)."""
    if isinstance(request, FitRequest):
        voxels = request.measured.reshape(-1, request.measured.shape[-1]).shape[0]
        samples = request.measured.shape[-1]
        # data, fit state, and (sequences off the fused path) the signal and
        # Jacobian one iteration forms, which a captured iteration keeps
        columns = len(request.unknown)
        return voxels * (samples * 8 * (2 + columns) + 8 * 40 + 4 * 8)
    sim = request.simulator
    held, sequence = sim._split({**sim.bound, **request.properties})
    shape = sim._batch(held)
    atoms = 1
    for size in shape:
        atoms *= int(size)
    packed = sim.packed(**sequence)
    return atoms * (packed.output_count * 8 * (1 + len(request.diff)) + 17 * 4) + packed.event_count * 64


class Server:
    """Serve simulate and fit requests on one GPU (see the module docstring).

    ``hbm_budget`` in bytes (default: half of the device memory free at
    start); ``streams`` HIP streams at the default priority (requests rotate
    over them: a request queued behind a long one on the same stream waits
    for it, so more streams than HIP's default 4 hardware queues pay off;
    ``GPU_MAX_HW_QUEUES`` set before the process first touches the GPU
    raises those); ``tick`` seconds the scheduler sleeps when it has nothing
    to do.
    """

    def __init__(self, device: str | torch.device = "cuda", *, hbm_budget: int | None = None, streams: int = 16,
                 tick: float = 2e-4, grid_blocks: int = 0, graphs: bool = False) -> None:
        if graphs:
            # Replaying captured fit iterations faulted the GPU under the
            # 60 s serving mix (illegal address, twice, the second time with
            # graph release deferred to the request's completion); the cause
            # is not found, so the path stays off.
            raise NotImplementedError("capturing fit iterations into HIP graphs is switched off")
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        _native.require_device(self.device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * 0.5)
        self.hbm_budget = int(hbm_budget)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(max(1, streams))]
        self._rotation = itertools.cycle(range(len(self.streams)))
        self.tick = tick
        self.grid_blocks = grid_blocks
        self.graphs = graphs
        self._incoming: queue.Queue = queue.Queue()
        self._waiting: list[_Ticket] = []
        self._active: list[_Ticket] = []
        self._finishing: list[_Ticket] = []
        self._retired: list[tuple[torch.cuda.Event, Any]] = []  # failed programs whose work may be in flight
        self._used = 0
        self.resident: dict[str, Any] = {}
        self._resident_bytes: dict[str, int] = {}
        self._lock = threading.Lock()
        self._stop = threading.Event()
        self._thread: threading.Thread | None = None
        self.latencies: list[float] = []
        self.host_seconds = {"admit": 0.0, "step": 0.0, "batched": 0.0, "unbatched": 0.0, "complete": 0.0}
        self.launches = 0
        self.batched_jobs = 0
        self.completed = 0
        self.started_at: float | None = None

    # -- public -------------------------------------------------------------

    def start(self) -> Server:
        if self._thread is None:
            self._stop.clear()
            self.started_at = time.perf_counter()
            self._thread = threading.Thread(target=self._loop, name="tsim-server", daemon=True)
            self._thread.start()
        return self

    def stop(self, wait: bool = True) -> None:
        self._stop.set()
        if self._thread is not None and wait:
            self._thread.join()
        self._thread = None

    def __enter__(self) -> Server:
        return self.start()

    def __exit__(self, *exc) -> None:
        self.stop()

    def submit(self, request: Any) -> Future:
        """Queue a request; the future resolves to its result."""
        future: Future = Future()
        size = _request_bytes(request)
        if size > self.hbm_budget:
            future.set_exception(MemoryError(f"request needs {size} bytes, over the {self.hbm_budget}-byte budget"))
            return future
        self._incoming.put(_Ticket(request, future, time.perf_counter(), size))
        return future

    def release(self, name: str) -> None:
        """Drop a resident result and return its bytes to the budget."""
        with self._lock:
            self.resident.pop(name, None)
            self._used -= self._resident_bytes.pop(name, 0)

    def stats(self) -> dict[str, float]:
        values = sorted(self.latencies)

        def pct(q):
            if not values:
                return float("nan")
            return values[min(len(values) - 1, int(round(q * (len(values) - 1))))]

        return {
            "completed": self.completed,
            "p50_s": pct(0.5),
            "p99_s": pct(0.99),
            "launches": self.launches,
            "batched_jobs": self.batched_jobs,
            "hbm_budget": self.hbm_budget,
            # host seconds the scheduler spent building, launching and retiring work
            "host_s": dict(self.host_seconds),
        }

    # -- the loop ------------------------------------------------------------

    def _loop(self) -> None:
        torch.cuda.set_device(self.device)
        while (not self._stop.is_set() or self._waiting or self._active or self._finishing or self._retired
               or not self._incoming.empty()):
            busy = False
            t0 = time.perf_counter()
            busy |= self._admit()
            t1 = time.perf_counter()
            busy |= self._step()
            t2 = time.perf_counter()
            busy |= self._complete()
            self.host_seconds["admit"] += t1 - t0
            self.host_seconds["step"] += t2 - t1
            self.host_seconds["complete"] += time.perf_counter() - t2
            if not busy:
                time.sleep(self.tick)

    def _admit(self) -> bool:
        while True:
            try:
                self._waiting.append(self._incoming.get_nowait())
            except queue.Empty:
                break
        admitted = False
        still = []
        for ticket in self._waiting:
            with self._lock:
                fits = self._used + ticket.bytes <= self.hbm_budget
                if fits:
                    self._used += ticket.bytes
            if not fits:
                still.append(ticket)
                continue
            ticket.admitted = time.perf_counter()
            ticket.stream = self.streams[next(self._rotation)]
            try:
                with torch.cuda.stream(ticket.stream):
                    if isinstance(ticket.request, FitRequest):
                        ticket.program = _FitProgram(ticket.request, self.device, graphs=self.graphs)
                    else:
                        ticket.program = _SimulateProgram(ticket.request, self.device)
            except Exception as error:  # a request that cannot be built fails alone
                self._fail(ticket, error)
                continue
            ticket.last = torch.cuda.Event()
            ticket.last.record(ticket.stream)
            self._active.append(ticket)
            admitted = True
        self._waiting = still
        return admitted

    def _fail(self, ticket: _Ticket, error: BaseException) -> None:
        with self._lock:
            self._used -= ticket.bytes
        if ticket.program is not None and ticket.stream is not None:
            # its queued work may still read what the program holds
            passed = torch.cuda.Event()
            passed.record(ticket.stream)
            self._retired.append((passed, ticket.program))
        ticket.future.set_exception(error)

    def _step(self) -> bool:
        if not self._active:
            return False
        groups: dict[tuple, list[tuple[_Ticket, BatchItem]]] = {}
        unbatched = []
        for ticket in self._active:
            try:
                with torch.cuda.stream(ticket.stream):
                    job = ticket.program.job()
                    if job is not None:
                        ticket.last = torch.cuda.Event()
                        ticket.last.record(ticket.stream)
            except Exception as error:
                self._fail(ticket, error)
                ticket.program = None
                continue
            if job is None:
                if not ticket.program.finished:
                    unbatched.append(ticket)
                continue
            mode, item = job
            key = (mode, *item.variant(), _tile_key(item))
            groups.setdefault(key, []).append((ticket, item))
        self._active = [t for t in self._active if t.program is not None]
        batched_from = time.perf_counter()
        for key, members in groups.items():
            stream = self.streams[next(self._rotation)]
            for ticket, _ in members:
                stream.wait_event(ticket.last)
            try:
                launch_batch([item for _, item in members], mode=key[0], stream=stream,
                             grid_blocks=self.grid_blocks)
                launched = torch.cuda.Event()
                launched.record(stream)
                for ticket, _ in members:
                    # the request's own work continues on its own stream
                    ticket.stream.wait_event(launched)
                    with torch.cuda.stream(ticket.stream):
                        ticket.program.after_job()
                        ticket.last = torch.cuda.Event()
                        ticket.last.record(ticket.stream)
            except Exception as error:
                for ticket, _ in members:
                    self._fail(ticket, error)
                    ticket.program = None
                continue
            self.launches += 1
            self.batched_jobs += len(members)
        self.host_seconds["batched"] += time.perf_counter() - batched_from
        unbatched_from = time.perf_counter()
        for ticket in unbatched:
            try:
                with torch.cuda.stream(ticket.stream):
                    ticket.program.run_unbatched()
                    ticket.last = torch.cuda.Event()
                    ticket.last.record(ticket.stream)
            except Exception as error:
                self._fail(ticket, error)
                ticket.program = None
        self.host_seconds["unbatched"] += time.perf_counter() - unbatched_from
        self._active = [t for t in self._active if t.program is not None]
        done = [t for t in self._active if t.program.finished]
        for ticket in done:
            ticket.done_event = ticket.last
            self._finishing.append(ticket)
        self._active = [t for t in self._active if not t.program.finished]
        return bool(groups or unbatched)

    def _complete(self) -> bool:
        self._retired = [(event, program) for event, program in self._retired if not event.query()]
        finished = [t for t in self._finishing if t.done_event.query()]
        if not finished:
            return False
        now = time.perf_counter()
        for ticket in finished:
            result = ticket.program.result
            request = ticket.request
            keep = isinstance(request, SimulateRequest) and request.resident
            with self._lock:
                if keep:
                    self.resident[request.resident] = result
                    self._resident_bytes[request.resident] = ticket.bytes
                else:
                    self._used -= ticket.bytes
            # the GPU has passed the request's last work: its graph may go
            release = getattr(ticket.program, "release", None)
            if release is not None:
                release()
            self.latencies.append(now - ticket.arrival)
            self.completed += 1
            ticket.future.set_result(result)
        self._finishing = [t for t in self._finishing if t not in finished]
        return True


def _tile_key(item: BatchItem) -> int:
    """Requests of one batch share a tile; group state counts into bands."""
    s = item.states
    for band in (4, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64):
        if s <= band:
            return band
    return s
