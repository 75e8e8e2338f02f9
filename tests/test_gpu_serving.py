"""The serving loop: continuous batching over the uneven-batch launcher."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

from tests import _cases

pytestmark = pytest.mark.gpu


def _mrf(pulses, states=10):
    from torchsim_amd.simulators import MRFSimulator

    return MRFSimulator(flip=torch.tensor(_cases.benchmark_flips(pulses), dtype=torch.float32), TR=10.0,
                        states=states)


def _simulate_request(atoms, pulses, seed, diff=("T1", "T2"), states=10):
    from torchsim_amd.serving import SimulateRequest

    t1, t2 = _cases.random_tissue(atoms, seed=seed)
    return SimulateRequest(_mrf(pulses, states), {"T1": torch.tensor(t1).cuda(), "T2": torch.tensor(t2).cuda()},
                           diff=diff)


def test_batched_launch_matches_single_launches_bitwise():
    from torchsim_amd.serving import BatchItem, launch_batch
    from torchsim_amd.sequence import _engine

    items, singles = [], []
    for k, (atoms, pulses, states) in enumerate([(37, 120, 10), (500, 300, 7), (1, 80, 10), (2049, 200, 9)]):
        seq = _mrf(pulses, states)
        t1, t2 = _cases.random_tissue(atoms, seed=40 + k)
        packed = seq.packed(**seq.bound)
        tissue = _engine.prepare_tissue({"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2)},
                                        torch.device("cuda", 0), ("t1_ms", "t2_ms"))
        sig = torch.empty((atoms, pulses), dtype=torch.complex64, device="cuda")
        jac = torch.empty((atoms, 2, pulses), dtype=torch.complex64, device="cuda")
        items.append(BatchItem(packed, tissue, states, ("t1_ms", "t2_ms"), 1j, signal=sig, jacobian=jac))
        singles.append(_engine.simulate_real(packed, tissue, states=states, columns=("t1_ms", "t2_ms"), factor=1j,
                                             tile=5))
    stream = torch.cuda.current_stream()
    launch_batch(items, mode=0, stream=stream, tile=5, grid_blocks=7)  # persistent: 7 workgroups walk the table
    torch.cuda.synchronize()
    for item, (sig, jac) in zip(items, singles):
        assert torch.equal(item.signal, sig)
        assert torch.equal(item.jacobian, jac)


def _complex_requests(diff=("T1", "T2", "M0")):
    """Requests that take the complex kernel, one per feature it adds: phase
    cycling with off-resonance, an exchanging pool with a chemical shift,
    diffusion through a crusher geometry with off-resonance, off-resonance
    alone; 6 to 13
    orders, 1 to 2049 atoms. The TRs ramp, so the winding is not uniform and
    the off-resonance stays in the states."""
    from torchsim_amd.serving import SimulateRequest
    from torchsim_amd.simulators import MRFSimulator

    rng = np.random.default_rng(50)
    requests = []
    for k, (atoms, pulses, states, kind) in enumerate(
            [(37, 90, 6, "cycled"), (500, 120, 10, "pool"), (1, 60, 13, "diffusion"), (2049, 150, 10, "b0")]):
        t1, t2 = _cases.random_tissue(atoms, seed=70 + k)
        props = {"T1": torch.tensor(t1).cuda(), "T2": torch.tensor(t2).cuda(),
                 "M0": torch.tensor(rng.uniform(0.5, 1.0, atoms), dtype=torch.float32).cuda()}
        sequence = {"flip": torch.tensor(_cases.benchmark_flips(pulses), dtype=torch.float32),
                    "TR": torch.linspace(10.0, 14.0, pulses)}
        if kind == "cycled":
            sequence["phases"] = torch.tensor((np.arange(pulses) ** 2 * 7.5) % 360.0, dtype=torch.float32)
        if kind in ("cycled", "b0", "diffusion"):
            props["B0"] = torch.tensor(rng.uniform(-20.0, 20.0, atoms), dtype=torch.float32).cuda()
        if kind == "pool":
            props.update(poolB_fraction=torch.tensor(rng.uniform(0.05, 0.3, atoms), dtype=torch.float32).cuda(),
                         poolB_exchange=torch.tensor(rng.uniform(5.0, 20.0, atoms), dtype=torch.float32).cuda(),
                         poolB_T1=400.0, poolB_T2=15.0, poolB_shift=35.0)
        if kind == "diffusion":
            props.update(D=torch.tensor(rng.uniform(0.5, 3.0, atoms), dtype=torch.float32).cuda(),
                         crusher_dephasing_rad=8 * np.pi, voxel_size_m=0.5e-3)
        requests.append(SimulateRequest(MRFSimulator(states=states, **sequence), props, diff=diff))
    return requests


def test_complex_batch_matches_single_launches_bitwise():
    """Single-pool, exchanging-pool, diffusing and phase-cycled requests of
    different sizes and order counts in ONE persistent launch of the complex
    kernel: every signal and Jacobian column equals its own launch bitwise,
    with and without Jacobian columns."""
    from torchsim_amd.serving import BatchItem, launch_batch
    from torchsim_amd.sequence import _complex, _engine

    for diff in (("T1", "T2", "M0"), (), "mixed"):
        items, singles = [], []
        if diff == "mixed":
            # one tangent set (none), stream counts 1 and 2 per atom in one batch:
            # the waves of a block stage different requests side by side
            plain, with_m0 = _complex_requests(()), _complex_requests(("M0",))
            requests = [plain[k] if k % 2 == 0 else with_m0[k] for k in range(len(plain))]
        else:
            requests = _complex_requests(diff)
        for request in requests:
            sim = request.simulator
            held, sequence = sim._split({**sim.bound, **request.properties})
            _, packed, _, tissue, states, columns, _, _, _ = sim._prepared(held, sequence, request.diff)
            assert not _engine.real_subspace(packed, tissue)[0]
            atoms, n = tissue.atoms, packed.output_count
            sig = torch.full((atoms, n), float("nan"), dtype=torch.complex64, device="cuda")
            jac = torch.full((atoms, len(columns), n), float("nan"), dtype=torch.complex64, device="cuda")
            items.append(BatchItem(packed, tissue, states, columns, sim.signal_factor, signal=sig, jacobian=jac,
                                   kernel="complex"))
            singles.append(_complex.simulate_complex(packed, tissue, states=states, columns=columns,
                                                     factor=sim.signal_factor))
        print(f"case: complex batch, columns {diff}", flush=True)
        assert len({item.variant() for item in items}) == 1
        launch_batch(items, mode=0, stream=torch.cuda.current_stream(), grid_blocks=7)
        torch.cuda.synchronize()
        for item, (sig, jac) in zip(items, singles):
            assert torch.equal(item.signal, sig)
            if item.columns:
                assert torch.equal(item.jacobian, jac)


def test_a_complex_batch_refuses_what_it_does_not_run():
    from torchsim_amd.serving import BatchItem, launch_batch
    from torchsim_amd.sequence import _engine

    request = _complex_requests()[0]
    sim = request.simulator
    held, sequence = sim._split({**sim.bound, **request.properties})
    _, packed, _, tissue, states, _, _, _, _ = sim._prepared(held, sequence, ())
    sig = torch.empty((tissue.atoms, packed.output_count), dtype=torch.complex64, device="cuda")
    item = BatchItem(packed, tissue, states, ("inversion_efficiency",), 1j, signal=sig, jacobian=sig,
                     kernel="complex")
    with pytest.raises(ValueError, match="no Jacobian column"):
        launch_batch([item], mode=0, stream=torch.cuda.current_stream())
    with pytest.raises(ValueError, match="mode 0"):
        launch_batch([dataclasses.replace(item, columns=())], mode=1, stream=torch.cuda.current_stream())
    real = dataclasses.replace(item, columns=(), kernel="real")
    with pytest.raises(ValueError, match="one kernel"):
        launch_batch([dataclasses.replace(item, columns=()), real], mode=0, stream=torch.cuda.current_stream())
    assert _engine is not None


def test_batched_launch_replays_from_a_hip_graph():
    """The uneven batch captured once into a HIP graph: every replay equals an
    eager launch bitwise, including after the tissue changed in place."""
    from torchsim_amd.serving import BatchItem, launch_batch
    from torchsim_amd.sequence import _engine

    items = []
    for k, (atoms, pulses, states) in enumerate([(300, 150, 10), (1025, 220, 8), (7, 90, 10)]):
        seq = _mrf(pulses, states)
        t1, t2 = _cases.random_tissue(atoms, seed=60 + k)
        tissue = _engine.prepare_tissue({"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2)},
                                        torch.device("cuda", 0), ("t1_ms", "t2_ms"))
        sig = torch.empty((atoms, pulses), dtype=torch.complex64, device="cuda")
        jac = torch.empty((atoms, 2, pulses), dtype=torch.complex64, device="cuda")
        items.append(BatchItem(seq.packed(**seq.bound), tissue, states, ("t1_ms", "t2_ms"), 1j, signal=sig,
                               jacobian=jac))

    def eager():
        launch_batch(items, mode=0, stream=torch.cuda.current_stream(), tile=5)
        torch.cuda.synchronize()
        return [(i.signal.clone(), i.jacobian.clone()) for i in items]

    def cleared():
        for i in items:
            i.signal.fill_(float("nan"))
            i.jacobian.fill_(float("nan"))

    print("case: eager warm-up", flush=True)
    eager()
    reference = eager()
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    print("case: capture", flush=True)
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            launch_batch(items, mode=0, stream=side, tile=5)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    for replay in range(3):
        print(f"case: replay {replay}", flush=True)
        cleared()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for item, (sig, jac) in zip(items, reference):
            assert torch.equal(item.signal, sig)
            assert torch.equal(item.jacobian, jac)
    # new tissue values, same buffers: the graph recomputes from them
    for item in items:
        item.tissue.buffers[0].mul_(1.1)
    changed = eager()
    assert not torch.equal(changed[0][0], reference[0][0])
    cleared()
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    for item, (sig, jac) in zip(items, changed):
        assert torch.equal(item.signal, sig)
        assert torch.equal(item.jacobian, jac)


def test_same_alone_and_in_the_mix():
    from torchsim_amd.serving import FitRequest, Server

    probe = _simulate_request(300, 250, seed=1)
    with Server() as alone:
        a_sig, a_jac = alone.submit(probe).result(timeout=120)
    seq = _mrf(200)
    t1, t2 = _cases.random_tissue(64, seed=7)
    clean = seq.simulate(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda())
    fit = FitRequest(seq, clean, bounds={"T1": (100.0, 5000.0), "T2": (10.0, 2000.0), "M0": (0.01, 2.0)},
                     initial={"T1": 1000.0, "T2": 100.0, "M0": 0.8}, iterations=6)
    with Server() as alone:
        f_alone = alone.submit(fit).result(timeout=120)
    with Server() as mixed:
        others = [mixed.submit(_simulate_request(n, p, seed=s)) for n, p, s in
                  [(1000, 400, 2), (17, 100, 3), (5000, 300, 4), (123, 600, 5)]]
        f_mix = mixed.submit(fit)
        mine = mixed.submit(probe)
        more = [mixed.submit(_simulate_request(n, p, seed=s, diff=())) for n, p, s in [(2000, 250, 6), (9, 50, 8)]]
        b_sig, b_jac = mine.result(timeout=120)
        f_mixed = f_mix.result(timeout=120)
        for f in others + more:
            f.result(timeout=120)
        stats = mixed.stats()
    assert torch.equal(a_sig, b_sig) and torch.equal(a_jac, b_jac)
    for name in ("T1", "T2", "M0"):
        assert torch.equal(f_alone[name], f_mixed[name]), name
    assert stats["completed"] == 8
    assert stats["by_kind"]["fit"]["completed"] == 1 and stats["by_kind"]["simulate"]["completed"] == 7
    assert stats["by_kind"]["fit"]["p99_s"] > 0.0
    assert stats["batched_jobs"] > stats["launches"]  # some launches carried several requests


@pytest.mark.skip(reason="graph replay of fit iterations is switched off: it faulted the GPU (illegal address) "
                         "under the 60 s serving mix, twice; the cause is not found")
def test_unbatched_fits_replayed_from_a_graph_equal_eager_iterations():
    """Fits off the fused path (an exchanging pool; FSE's recovery) run their
    iterations from a captured HIP graph: the maps equal those of eager
    iterations bitwise, alone and with other requests in flight."""
    from torchsim_amd.serving import FitRequest, Server
    from torchsim_amd.simulators import FSESimulator

    rng = np.random.default_rng(21)
    voxels = 96
    known = {"pool_b_fraction": torch.tensor(rng.uniform(0.05, 0.3, voxels), dtype=torch.float32).cuda(),
             "pool_b_exchange": torch.tensor(rng.uniform(5, 20, voxels), dtype=torch.float32).cuda(),
             "T1_pool_b": 400.0, "T2_pool_b": 15.0}
    t1 = torch.tensor(rng.uniform(400, 2000, voxels), dtype=torch.float32).cuda()
    t2 = torch.tensor(rng.uniform(40, 200, voxels), dtype=torch.float32).cuda()
    mrf = _mrf(160)
    pooled = mrf.simulate(T1=t1, T2=t2, **known)
    fse = FSESimulator(flip=torch.full((48,), 150.0), ESP=5.0, TR=3000.0)
    echoes = fse.simulate(T1=t1, T2=t2)
    bounds = {"T1": (100.0, 5000.0), "T2": (10.0, 2000.0), "M0": (0.01, 2.0)}
    start = {"T1": 1000.0, "T2": 100.0, "M0": 0.8}
    requests = [FitRequest(mrf, pooled, bounds=bounds, initial=start, known=known, iterations=7),
                FitRequest(fse, echoes, bounds=bounds, initial=start, iterations=7)]
    results = {}
    for graphs in (False, True):
        with Server(graphs=graphs) as server:
            results[graphs] = [server.submit(r).result(timeout=120) for r in requests]
    with Server(graphs=True) as mixed:
        noise = [mixed.submit(_simulate_request(n, p, seed=s)) for n, p, s in [(3000, 300, 31), (50, 120, 32)]]
        futures = [mixed.submit(r) for r in requests]
        mixed_maps = [f.result(timeout=120) for f in futures]
        for f in noise:
            f.result(timeout=120)
    for eager, replayed, in_mix in zip(results[False], results[True], mixed_maps):
        for name in ("T1", "T2", "M0"):
            assert torch.equal(eager[name], replayed[name]), name
            assert torch.equal(eager[name], in_mix[name]), name


def test_served_results_equal_the_api_and_fits_recover():
    from torchsim_amd.serving import FitRequest, Server

    request = _simulate_request(700, 300, seed=12)
    direct = request.simulator.jacobian(("T1", "T2"), **request.properties)
    seq = _mrf(300)
    t1 = np.linspace(400, 2000, 50).astype(np.float32)
    t2 = np.linspace(40, 200, 50).astype(np.float32)
    clean = seq.simulate(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda())
    fit = FitRequest(seq, clean, bounds={"T1": (100.0, 5000.0), "T2": (10.0, 2000.0), "M0": (0.01, 2.0)},
                     initial={"T1": 1000.0, "T2": 100.0, "M0": 0.8}, iterations=25)
    with Server() as server:
        served = server.submit(request).result(timeout=120)
        maps = server.submit(fit).result(timeout=120)
    assert torch.equal(served[0], direct[0]) and torch.equal(served[1], direct[1])
    assert float(((maps["T1"].cpu() - torch.tensor(t1)).abs() / torch.tensor(t1)).max()) < 1e-3
    assert float(((maps["T2"].cpu() - torch.tensor(t2)).abs() / torch.tensor(t2)).max()) < 1e-3


def test_complex_and_closed_form_requests_same_alone_and_in_the_mix():
    """bSSFP (closed form), exchanging-pool, diffusing, phase-cycled and
    off-resonant simulate requests and a two-pool fit, each served alone and
    all in one mix: the same bits, equal to the simulators' own calls; the
    complex-kernel requests of one tangent set share launches."""
    from torchsim_amd.serving import FitRequest, Server, SimulateRequest
    from torchsim_amd.simulators import bSSFPSimulator

    rng = np.random.default_rng(90)
    t1 = torch.tensor(rng.uniform(300, 2500, (6, 5)), dtype=torch.float32).cuda()
    t2 = torch.tensor(rng.uniform(20, 200, (6, 5)), dtype=torch.float32).cuda()
    bssfp = bSSFPSimulator(flip=torch.linspace(5.0, 60.0, 40), TR=5.0, TE=2.5)
    requests = [SimulateRequest(bssfp, {"T1": t1, "T2": t2}, diff=("T1", "T2")),
                *_complex_requests(("T1", "T2")), *_complex_requests(())]
    mrf = _mrf(80)
    voxels = 24
    known = {"poolB_fraction": torch.tensor(rng.uniform(0.05, 0.3, voxels), dtype=torch.float32).cuda(),
             "poolB_exchange": torch.tensor(rng.uniform(5, 20, voxels), dtype=torch.float32).cuda(),
             "poolB_T1": 400.0, "poolB_T2": 15.0}
    pooled = mrf.simulate(T1=torch.tensor(rng.uniform(400, 2000, voxels), dtype=torch.float32).cuda(),
                          T2=torch.tensor(rng.uniform(40, 200, voxels), dtype=torch.float32).cuda(), **known)
    fit = FitRequest(mrf, pooled, bounds={"T1": (100.0, 5000.0), "T2": (10.0, 2000.0), "M0": (0.01, 2.0)},
                     initial={"T1": 1000.0, "T2": 100.0, "M0": 0.8}, known=known, iterations=4)

    def direct(request):
        sim = request.simulator
        if request.diff:
            return sim.jacobian(request.diff, **request.properties)
        return sim.simulate(**request.properties)

    def same(a, b):
        if isinstance(a, tuple):
            return all(torch.equal(x, y) for x, y in zip(a, b))
        if isinstance(a, dict):
            return all(torch.equal(a[k], b[k]) for k in ("T1", "T2", "M0"))
        return torch.equal(a, b)

    alone = []
    for request in [*requests, fit]:
        with Server() as server:
            alone.append(server.submit(request).result(timeout=120))
    print("case: served alone", flush=True)
    with Server() as mixed:
        futures = [mixed.submit(request) for request in [*requests, fit]]
        together = [f.result(timeout=120) for f in futures]
        stats = mixed.stats()
    print("case: served in the mix", stats, flush=True)
    for k, request in enumerate(requests):
        assert same(alone[k], together[k]), k
        assert same(direct(request), together[k]), k
    assert same(alone[-1], together[-1])
    assert stats["completed"] == len(requests) + 1
    assert stats["batched_jobs"] > stats["launches"]


@pytest.mark.skip(reason="fits off the fused path as jobs of the complex kernel's batch are switched off: with "
                         "them the 60 s serving mix faulted the GPU (memory access) although this test passed; the "
                         "cause is not found, so those fits run ordinary iterations")
def test_a_two_pool_fit_in_the_complex_batch_equals_its_ordinary_iterations():
    """A fit off the fused path (an exchanging pool) runs each iteration's
    signal + Jacobian as a job of the complex kernel's batch, beside simulate
    requests of the same tangent set; its maps are the bits of the same
    Levenberg-Marquardt iterations run on the ordinary path."""
    from torchsim_amd.serving import FitRequest, Server
    from torchsim_amd.serving._server import _FitProgram

    rng = np.random.default_rng(91)
    mrf = _mrf(120)
    voxels = 40
    known = {"poolB_fraction": torch.tensor(rng.uniform(0.05, 0.3, voxels), dtype=torch.float32).cuda(),
             "poolB_exchange": torch.tensor(rng.uniform(5, 20, voxels), dtype=torch.float32).cuda(),
             "poolB_T1": 400.0, "poolB_T2": 15.0}
    truth_t1 = torch.tensor(rng.uniform(400, 2000, voxels), dtype=torch.float32).cuda()
    truth_t2 = torch.tensor(rng.uniform(40, 200, voxels), dtype=torch.float32).cuda()
    pooled = mrf.simulate(T1=truth_t1, T2=truth_t2, **known)
    fit = FitRequest(mrf, pooled, bounds={"T1": (100.0, 5000.0), "T2": (10.0, 2000.0), "M0": (0.01, 2.0)},
                     initial={"T1": 1000.0, "T2": 100.0, "M0": 0.8}, known=known, iterations=6)
    ordinary = _FitProgram(fit, torch.device("cuda", 0))
    assert ordinary.complex_batch and not ordinary.fused
    ordinary.complex_batch = False
    while not ordinary.finished:
        ordinary.run_unbatched()
    expect = ordinary.result
    with Server() as server:
        others = [server.submit(request) for request in _complex_requests(("T1", "T2"))]
        maps = server.submit(fit).result(timeout=120)
        for f in others:
            f.result(timeout=120)
        stats = server.stats()
    for name in ("T1", "T2", "M0"):
        assert torch.equal(maps[name], expect[name]), name
    assert stats["batched_jobs"] > stats["launches"]
    # and the fit moved toward the truth
    assert float(((maps["T1"] - truth_t1).abs() / truth_t1).median()) < 0.05


def test_unbatched_sequences_are_served_too():
    """FSE (complex kernel): a simulate request takes the complex batch, a
    fit the ordinary path."""
    from torchsim_amd.serving import FitRequest, Server, SimulateRequest
    from torchsim_amd.simulators import FSESimulator

    fse = FSESimulator(flip=torch.full((64,), 150.0), ESP=5.0, TR=3000.0)
    t1 = np.linspace(500, 2000, 20).astype(np.float32)
    t2 = np.linspace(40, 150, 20).astype(np.float32)
    props = {"T1": torch.tensor(t1).cuda(), "T2": torch.tensor(t2).cuda()}
    direct = fse.simulate(**props)
    with Server() as server:
        served = server.submit(SimulateRequest(fse, props)).result(timeout=120)
        maps = server.submit(FitRequest(fse, direct, bounds={"T1": (100.0, 5000.0), "T2": (10.0, 2000.0),
                                                             "M0": (0.01, 2.0)},
                                        initial={"T1": 1000.0, "T2": 80.0, "M0": 0.9}, iterations=30)).result(120)
    assert torch.equal(served, direct)
    assert float(((maps["T2"].cpu() - torch.tensor(t2)).abs() / torch.tensor(t2)).max()) < 1e-2


def test_admission_by_budget_and_refusal():
    from torchsim_amd.serving import Server

    small = _simulate_request(4000, 300, seed=20)
    from torchsim_amd.serving._server import _request_bytes

    need = _request_bytes(small)
    with Server(hbm_budget=int(need * 1.5)) as server:
        futures = [server.submit(_simulate_request(4000, 300, seed=20 + k)) for k in range(6)]
        for f in futures:
            f.result(timeout=120)
        huge = server.submit(_simulate_request(200000, 1000, seed=9))
        with pytest.raises(MemoryError):
            huge.result(timeout=10)
        assert server.stats()["completed"] == 6
