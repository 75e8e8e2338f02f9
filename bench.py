"""Throughput of the fused EPG signal + Jacobian on one MI355X.

Workload: the MRF dictionary -- 880 pulses (5->60 deg over 300, 60->2 over
300, 280 at 2 deg), TR 10 ms, an inversion at TI 0; T1 log-spaced 100-5000 ms
(100) x T2 log-spaced 10-2000 ms (80) with T2 < T1, x B1 0.60:0.05:1.40 (17)
= 106,182 atoms, tiled 12 times to 1,274,184 atoms; 20 EPG states; signal and
Jacobian w.r.t. (T1, T2, B1) as complex64. One step is one
``MRFSimulator.jacobian(("T1", "T2", "B1"), ...)`` call over the whole
dictionary with the tissue already resident in HBM.

Prints one JSON line: whole-job atom-pulses/s, a roofline object for the
dominant kernel (live HIP-event timing; FP32 VALU bound), and the CPU baseline
(the reference's own CPU kernel built from source, or the FP64 port) timed on
a bounded sample on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PULSES = 880
STATES = 20
NAMES = ("T1", "T2", "B1")
TILES = 12
#: Algorithmic FP32 flops per atom-pulse, real subspace, S states, P tangents:
#: S * (18 + 42 P) (FMA = 2).
FLOPS_PER_ATOM_PULSE = STATES * (18 + 42 * len(NAMES))
#: Bytes written per atom-pulse: complex64 signal + P complex64 columns.
BYTES_PER_ATOM_PULSE = 8 * (1 + len(NAMES))
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def flips() -> np.ndarray:
    return np.concatenate([np.linspace(5, 60, 300), np.linspace(60, 2, 300), np.full(280, 2.0)])


def dictionary(tiles: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    t1 = np.geomspace(100.0, 5000.0, 100)
    t2 = np.geomspace(10.0, 2000.0, 80)
    b1 = np.round(np.arange(0.60, 1.40 + 1e-9, 0.05), 2)
    T1, T2 = np.meshgrid(t1, t2, indexing="ij")
    keep = T2 < T1
    pt1, pt2 = T1[keep], T2[keep]
    T1b = np.repeat(pt1[:, None], b1.size, 1).reshape(-1)
    T2b = np.repeat(pt2[:, None], b1.size, 1).reshape(-1)
    B1b = np.tile(b1, pt1.size)
    return tuple(np.tile(x, tiles).astype(np.float32) for x in (T1b, T2b, B1b))


def cpu_baseline(budget_s: float = 12.0) -> dict:
    """The reference's CPU kernel (or the FP64 port) on a sample of the same workload.

    What the reference does for a 3-property Jacobian: one forward and one JVP
    pass per property (reference model/_signal.py:230-238), all host cores.
    """
    import torch

    import oracle
    from torchsim_amd.simulators import MRFSimulator

    packed = MRFSimulator().packed(flip=torch.tensor(flips(), dtype=torch.float32), TR=10.0)
    events = oracle.events_of(packed)
    T1, T2, B1 = dictionary(1)
    cores = len(os.sched_getaffinity(0))
    module = None
    try:
        module = oracle.reference_module()
    except Exception:  # pragma: no cover - a broken build is reported as absent
        module = None
    rng = np.random.default_rng(0)
    atoms = 256
    while True:
        pick = rng.choice(T1.size, atoms, replace=False)
        tissue = {"t1": T1[pick], "t2": T2[pick], "b1": B1[pick]}
        start = time.perf_counter()
        if module is not None:
            for name in ("t1", "t2", "b1"):
                oracle.reference_real(events, tissue, states=STATES, outputs=PULSES, features=8, threads=cores)
                oracle.reference_real(
                    events, tissue, states=STATES, outputs=PULSES, features=8, threads=cores,
                    tangent={name: np.ones(atoms, dtype=np.float32)},
                )
        else:
            oracle.real(
                events, t1=tissue["t1"], t2=tissue["t2"], b1=tissue["b1"], states=STATES, outputs=PULSES,
                tangents=[{"t1": 1.0}, {"t2": 1.0}, {"b1": 1.0}],
            )
        elapsed = time.perf_counter() - start
        if elapsed > budget_s / 4 or atoms >= T1.size:
            break
        atoms = min(T1.size, atoms * 4)
    return {
        "value": atoms * PULSES / elapsed,
        "unit": "atom*pulses/s",
        "cores": cores if module is not None else 1,
        "kind": "reference" if module is not None else "port",
        "sample": f"{atoms} dictionary atoms x {PULSES} pulses, S={STATES}, signal + d/d(T1,T2,B1): "
        + ("3 x (simulate + simulate_jvp) of the reference _epg_cpu built from source" if module is not None
           else "FP64 C port, one pass with 3 tangents"),
        "seconds": elapsed,
    }


def torch_per_op_baseline(atoms: int = 65536) -> dict:
    """The plain per-op torch-ROCm EPG path on the GPU: the same recursion
    (real subspace, value + T1/T2/B1 tangents) as whole-array torch
    operations, one group of launches per event -- the way a torch program
    without a fused kernel runs it (the reference's own test oracle,
    tests/utils/packed_reference.py, is this pattern in complex arithmetic)."""
    import torch

    from torchsim_amd.simulators import MRFSimulator

    device = torch.device("cuda", 0)
    packed = MRFSimulator().packed(flip=torch.tensor(flips(), dtype=torch.float32), TR=10.0)
    T1, T2, B1 = dictionary(1)
    pick = np.random.default_rng(0).choice(T1.size, atoms, replace=False)
    t1 = torch.tensor(T1[pick], device=device)
    t2 = torch.tensor(T2[pick], device=device)
    b1 = torch.tensor(B1[pick], device=device)
    events = [(float(d), int(k), float(f), int(a), int(o)) for d, k, f, a, o in zip(
        packed.duration.tolist(), packed.kind.tolist(), packed.flip.tolist(), packed.action.tolist(),
        packed.output_index.tolist())]

    def run():
        S = STATES
        a = torch.zeros((4, atoms, S), device=device)  # value, dT1, dT2, dB1
        b = torch.zeros_like(a)
        c = torch.zeros_like(a)
        c[0, :, 0] = 1.0
        out = torch.empty((atoms, 4, PULSES), device=device)
        r1, r2 = 1000.0 / t1, 1000.0 / t2

        def shift(x, y):
            y2 = torch.cat((y[..., 1:], torch.zeros_like(y[..., :1])), -1)
            x2 = torch.cat((-y2[..., :1], x[..., :-1]), -1)
            return x2, y2

        for dt, kind, flip, act, o in events:
            if dt != 0.0:
                e1 = torch.exp(-r1 * dt)[:, None]
                e2 = torch.exp(-r2 * dt)[:, None]
                de1 = e1 * (1000.0 * dt / t1**2)[:, None]
                de2 = e2 * (1000.0 * dt / t2**2)[:, None]
                a0, b0, c0 = a[0].clone(), b[0].clone(), c[0].clone()
                a, b, c = a * e2, b * e2, c * e1
                a[2] += a0 * de2
                b[2] += b0 * de2
                c[1] += c0 * de1
                c[0, :, 0] += 1 - e1[:, 0]
                c[1, :, 0] -= de1[:, 0]
            if act & 1:
                a, b = shift(a, b)
            if kind == 1 and not act & 4:
                alpha = flip * b1
                cs, sn = torch.cos(alpha)[:, None], torch.sin(alpha)[:, None]
                ch, sh, hs = 0.5 * (1 + cs), 0.5 * (1 - cs), 0.5 * sn
                p0, m0, z0 = a[0], b[0], c[0]
                na = ch * a + sh * b - sn * c
                nb = sh * a + ch * b + sn * c
                nc = hs * (a - b) + cs * c
                dal = flip
                na[3] += (-hs * (p0 - m0) - cs * z0) * dal
                nb[3] -= (-hs * (p0 - m0) - cs * z0) * dal
                nc[3] += (0.5 * cs * (p0 - m0) - sn * z0) * dal
                a, b, c = na, nb, nc
            elif kind == 1:
                c = -c
            elif kind == 2 and act & 32:
                out[:, :, o] = -a[:, :, 0].transpose(0, 1)
            if act & 2:
                a, b = shift(a, b)
            if act & 8:
                a, b = torch.zeros_like(a), torch.zeros_like(b)
            elif act & 16:
                a, b = shift(a, b)
        return out

    run()
    torch.cuda.synchronize()
    start = time.perf_counter()
    run()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    return {"value": atoms * PULSES / elapsed, "unit": "atom*pulses/s", "atoms": atoms, "seconds": elapsed,
            "sample": f"{atoms} dictionary atoms x {PULSES} pulses, S={STATES}, value + T1/T2/B1 tangents, "
                      "whole-array torch ops per event on the GPU (FP32)"}


def random_tissue(atoms: int, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """T1 ~ U[200, 3000] ms, T2 ~ U[20, 300] ms, T2 redrawn until T2 < T1
    (torch generator, seed fixed): the design and CPU-baseline tissue."""
    import torch

    generator = torch.Generator().manual_seed(seed)
    t1 = torch.empty(atoms, dtype=torch.float64).uniform_(200.0, 3000.0, generator=generator)
    t2 = torch.empty(atoms, dtype=torch.float64).uniform_(20.0, 300.0, generator=generator)
    bad = t2 >= t1
    while bool(bad.any()):
        t2[bad] = torch.empty(int(bad.sum()), dtype=torch.float64).uniform_(20.0, 300.0, generator=generator)
        bad = t2 >= t1
    return t1.numpy().astype(np.float32), t2.numpy().astype(np.float32)


def phantom(shape=(256, 256, 8), *, classes=6, seed=0):
    """Piecewise-smooth T1/T2/M0 maps: ``classes`` random ellipsoidal tissue
    classes (the first fills the field of view) with T1 300-2500 ms, T2
    30-250 ms, M0 0.5-1.0, each map times (1 + 5% Gaussian-smoothed noise)."""
    import torch

    generator = torch.Generator().manual_seed(seed)

    def draw(low, high, n):
        return low + (high - low) * torch.rand(n, generator=generator, dtype=torch.float64)

    t1c, t2c, m0c = draw(300, 2500, classes), draw(30, 250, classes), draw(0.5, 1.0, classes)
    axes = [torch.linspace(-1, 1, n, dtype=torch.float64) for n in shape]
    grid = torch.stack(torch.meshgrid(*axes, indexing="ij"), -1)
    label = torch.zeros(shape, dtype=torch.int64)
    for k in range(1, classes):
        center = draw(-0.5, 0.5, 3)
        radii = draw(0.15, 0.5, 3)
        label[(((grid - center) / radii) ** 2).sum(-1) <= 1.0] = k
    kernel = torch.exp(-0.5 * (torch.arange(-6, 7, dtype=torch.float64) / 2.0) ** 2)
    kernel = kernel / kernel.sum()
    maps = []
    for values in (t1c, t2c, m0c):
        base = values[label]
        smooth = torch.randn(shape, generator=generator, dtype=torch.float64)[None, None]
        for axis in range(3):
            view = [1, 1, 1]
            view[axis] = kernel.numel()
            pad = [0, 0, 0, 0, 0, 0]
            pad[2 * (2 - axis)] = pad[2 * (2 - axis) + 1] = 6
            smooth = torch.nn.functional.conv3d(torch.nn.functional.pad(smooth, pad, mode="replicate"),
                                                kernel.reshape(1, 1, *view))
        smooth = smooth[0, 0]
        maps.append((base * (1.0 + 0.05 * smooth / smooth.abs().max())).to(torch.float32))
    return tuple(maps)


def design_workload(steps: int, warmup: int) -> dict:
    """Sequence design: 1,024 atoms, 880 flips U[2,70] deg and TRs U[8,15] ms,
    loss sum tr(F^-1(T1,T2,M0)), gradient w.r.t. every flip and TR through
    the checkpointed adjoint (a checkpoint every 64 pulses)."""
    import torch

    from torchsim_amd.optim import crlb
    from torchsim_amd.simulators import MRFSimulator

    rng = np.random.default_rng(1)
    flips_deg = rng.uniform(2.0, 70.0, PULSES)
    trs = rng.uniform(8.0, 15.0, PULSES)
    t1, t2 = random_tissue(1024, seed=0)
    T1, T2 = torch.tensor(t1).cuda(), torch.tensor(t2).cuda()
    M0 = torch.ones(1024).cuda()
    seq = MRFSimulator(states=10, checkpoint_pulses=64)

    def step():
        flip = torch.tensor(flips_deg, dtype=torch.float32, requires_grad=True)
        tr = torch.tensor(trs, dtype=torch.float32, requires_grad=True)
        _, jac = seq.jacobian(("T1", "T2", "M0"), flip=flip, TR=tr, T1=T1, T2=T2, M0=M0)
        loss = crlb(jac).sum()
        loss.backward()
        return flip.grad, tr.grad

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - start) * 1e3 / steps
    return {"metric": "sequence-design gradient steps/s", "value": 1e3 / ms, "unit": "steps/s", "ms_per_step": ms,
            "config": {"workload": "crlb(T1,T2,M0) gradient w.r.t. 880 flips + 880 TRs", "atoms": 1024,
                       "pulses": PULSES, "states": 10, "checkpoint_pulses": 64},
            "atom_pulses_per_s": 1024 * PULSES / (ms * 1e-3)}


def fit_workload(iterations: int = 10) -> dict:
    """Gauss-Newton mapping of a 256x256x8 phantom (524,288 voxels) from
    880-frame MRF at SNR 50: 10 Levenberg-Marquardt iterations of T1, T2, M0."""
    import torch

    from torchsim_amd.estimators import NonlinearLeastSquares
    from torchsim_amd.recon import GaussNewton, TrustRegion
    from torchsim_amd.simulators import MRFSimulator

    t1, t2, m0 = (x.cuda() for x in phantom((256, 256, 8), seed=0))
    seq = MRFSimulator(flip=torch.tensor(flips(), dtype=torch.float32), TR=10.0, states=10)
    clean = seq.simulate(T1=t1, T2=t2, M0=m0)
    generator = torch.Generator(device="cuda").manual_seed(0)
    sigma = float(clean.abs().pow(2).mean().sqrt()) / 50.0
    noise = torch.complex(torch.randn(clean.shape, generator=generator, device="cuda"),
                          torch.randn(clean.shape, generator=generator, device="cuda")) * (sigma / np.sqrt(2))
    measured = clean + noise
    del clean, noise
    fit = NonlinearLeastSquares(seq, bounds={"T1": (100.0, 5000.0), "T2": (10.0, 2000.0), "M0": (0.01, 2.0)},
                                loop=GaussNewton(TrustRegion(), max_iterations=iterations))
    fit.fit(T1=(300.0, 2500.0), T2=(30.0, 250.0), M0=(0.5, 1.0), seed=0)
    # one untimed fit: code objects load and the state tile is measured once
    fit.map(measured)
    torch.cuda.synchronize()
    start = time.perf_counter()
    maps = fit.map(measured)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - start
    quantiles = torch.tensor([0.5, 0.9, 0.99], device="cuda", dtype=torch.float64)
    errors = {}
    for name, truth in (("T1", t1), ("T2", t2), ("M0", m0)):
        relative = ((maps[name] - truth).abs() / truth).double().reshape(-1)
        # (torch.quantile takes at most 2^24 values: the phantom's 524,288 fit)
        errors[name] = dict(zip(("p50", "p90", "p99"), torch.quantile(relative, quantiles).tolist()))
    voxels = t1.numel()
    return {"metric": "Gauss-Newton voxels/s (10 iterations)", "value": voxels / seconds, "unit": "voxels/s",
            "seconds": seconds, "iterations": fit.iterations, "unconverged": fit.unconverged,
            "relative_error_quantiles": errors,
            "config": {"workload": "256x256x8 phantom, 880-frame MRF, SNR 50, LM over T1/T2/M0", "voxels": voxels,
                       "states": 10}}


def serve_workload(duration_s: float = 60.0, rate: float = 50.0, seed: int = 0) -> dict:
    """Poisson arrivals (rate/s for duration_s): half simulate, half fit;
    voxels log-uniform 1k-200k; bSSFP 200-500 pulses, FSE 64-256 echoes or
    MRF 500-1000 pulses; a third with a two-pool exchange model, a sixth
    diffusion-weighted. A fit request's data is what a scanner would hand
    over: each voxel one of 256 tissue classes of its protocol plus noise.
    Reports p50/p99 latency (submission to result) and throughput."""
    import torch

    from torchsim_amd.serving import FitRequest, Server, SimulateRequest
    from torchsim_amd.simulators import FSESimulator, MRFSimulator, TrueFISPSimulator

    rng = np.random.default_rng(seed)
    count = int(duration_s * rate)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    geometry = dict(crusher_dephasing_rad=8 * np.pi, voxel_size_m=0.5e-3)
    plans = []
    for k in range(count):
        voxels = int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        family = rng.integers(3)
        pooled = rng.uniform() < 1 / 3
        diffusing = (not pooled) and rng.uniform() < 0.25  # one sixth of all requests
        if family == 0:
            sim = TrueFISPSimulator(flip=float(rng.uniform(20, 60)), TR=4.5, nshots=int(rng.integers(200, 501)),
                                    states=4 if diffusing else 1, **(geometry if diffusing else {}))
        elif family == 1:
            echoes = int(rng.integers(64, 257))
            sim = FSESimulator(flip=torch.full((echoes,), float(rng.uniform(120, 180))), ESP=5.0, TR=3000.0,
                               **(geometry if diffusing else {}))
        else:
            pulses = int(rng.integers(500, 1001))
            train = np.concatenate([np.linspace(5, 60, pulses // 2), np.linspace(60, 2, pulses - pulses // 2)])
            sim = MRFSimulator(flip=torch.tensor(train, dtype=torch.float32), TR=10.0, states=10,
                               **(geometry if diffusing else {}))
        plans.append((arrivals[k], voxels, family, pooled, diffusing, sim, rng.uniform() < 0.5,
                      int(rng.integers(1 << 30))))

    classes = 256  # tissue classes the measured data of a fit request is drawn from

    def tissue(voxels, pooled, diffusing, draw_seed):
        g = np.random.default_rng(draw_seed)
        props = {"T1": torch.tensor(g.uniform(300, 2500, voxels), dtype=torch.float32).cuda(),
                 "T2": torch.tensor(g.uniform(30, 250, voxels), dtype=torch.float32).cuda()}
        if pooled:
            props.update(pool_b_fraction=torch.tensor(g.uniform(0.05, 0.3, voxels), dtype=torch.float32).cuda(),
                         pool_b_exchange=torch.tensor(g.uniform(5, 20, voxels), dtype=torch.float32).cuda(),
                         T1_pool_b=400.0, T2_pool_b=15.0)
        if diffusing:
            props["D"] = torch.tensor(g.uniform(0.5, 3.0, voxels), dtype=torch.float32).cuda()
        return props

    def measured(sim, voxels, pooled, diffusing, draw_seed):
        """What a scanner hands a fit request: every voxel one of 256 tissue
        classes of the protocol (the classes simulated, the voxels gathered)
        plus complex noise at SNR 50, and the known per-voxel properties."""
        table = tissue(classes, pooled, diffusing, draw_seed)
        signals = sim.simulate(**table)
        g = torch.Generator(device="cuda").manual_seed(draw_seed)
        pick = torch.randint(0, classes, (voxels,), device="cuda", generator=g)
        data = signals[pick]
        sigma = signals.abs().max() / 50.0
        data = data + sigma * torch.randn(data.shape, dtype=data.dtype, device="cuda", generator=g)
        known = {k: (v[pick] if isinstance(v, torch.Tensor) else v) for k, v in table.items() if k not in ("T1", "T2")}
        return data, known

    bounds = {"T1": (100.0, 5000.0), "T2": (10.0, 2000.0), "M0": (0.01, 2.0)}
    client = torch.cuda.Stream()
    outcome = {"errors": 0, "done": 0}
    lock = __import__("threading").Lock()

    def finished(future):
        with lock:
            outcome["done"] += 1
            if future.exception() is not None:
                outcome["errors"] += 1

    with Server() as server:
        start = time.perf_counter()
        for when, voxels, family, pooled, diffusing, sim, is_fit, draw_seed in plans:
            delay = start + when - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            # the client builds each request as it arrives (its own stream)
            with torch.cuda.stream(client):
                if is_fit:
                    data, known = measured(sim, voxels, pooled, diffusing, draw_seed)
                    request = FitRequest(sim, data, bounds=bounds, known=known,
                                         initial={"T1": 1000.0, "T2": 100.0, "M0": 0.8}, iterations=10)
                else:
                    request = SimulateRequest(sim, tissue(voxels, pooled, diffusing, draw_seed), diff=("T1", "T2"))
            client.synchronize()
            server.submit(request).add_done_callback(finished)
        submitted = time.perf_counter() - start
        while True:
            with lock:
                if outcome["done"] >= count:
                    break
            time.sleep(0.01)
        elapsed = time.perf_counter() - start
        stats = server.stats()
    return {"metric": "served requests/s with p50/p99 latency", "value": count / elapsed, "unit": "req/s",
            "p50_s": stats["p50_s"], "p99_s": stats["p99_s"], "by_kind": stats["by_kind"],
            "completed": stats["completed"], "errors": outcome["errors"], "launches": stats["launches"], "batched_jobs": stats["batched_jobs"],
            "seconds": elapsed, "submission_s": submitted, "server_host_s": stats.get("host_s"),
            "data": "fit data: 256 tissue classes per protocol, simulated, gathered to the voxels, SNR 50 noise",
            "config": {"workload": "Poisson mix", "rate_per_s": rate, "duration_s": duration_s, "requests": count}}


FP16_MFMA_PEAK_TFLOPS = 2500.0


#: Algorithmic flops per configuration order per pulse of the exchanging
#: pool with tangents along T1 and T2 (complex multiply 6, complex add 2,
#: real-times-complex 2): relaxation of the value (F+ and F- through the 2x2
#: complex transverse exponential, 28 each; Z through the real longitudinal
#: one, 12) = 68; both tangents relaxed the same way (136) plus their sources,
#: dX F+ and conj(dX) F- added for T2 (64) and dY Z added for T1 (16) = 216;
#: the RF T-matrix on value and two tangents of both pools (3 x 2 x 54) = 324.
TWO_POOL_FLOPS_PER_ORDER = 68 + 216 + 324


def complex_workload(steps: int = 5, warmup: int = 2, atoms: int = 65536, pulses: int = 750) -> dict:
    """The complex kernel at a serving shape: a two-pool MRF protocol (fast
    fraction, exchange and pool-b relaxation per voxel), 65,536 atoms x 750
    pulses, 10 states, signal + d/d(T1, T2) in one launch."""
    import torch

    from torchsim_amd.simulators import MRFSimulator

    g = np.random.default_rng(0)
    device = torch.device("cuda", 0)
    train = np.concatenate([np.linspace(5, 60, pulses // 2), np.linspace(60, 2, pulses - pulses // 2)])
    sim = MRFSimulator(flip=torch.tensor(train, dtype=torch.float32), TR=10.0, states=10)
    props = {"T1": torch.tensor(g.uniform(300, 2500, atoms), dtype=torch.float32, device=device),
             "T2": torch.tensor(g.uniform(30, 250, atoms), dtype=torch.float32, device=device),
             "pool_b_fraction": torch.tensor(g.uniform(0.05, 0.3, atoms), dtype=torch.float32, device=device),
             "pool_b_exchange": torch.tensor(g.uniform(5, 20, atoms), dtype=torch.float32, device=device),
             "T1_pool_b": 400.0, "T2_pool_b": 15.0}
    for _ in range(warmup):
        sim.jacobian(("T1", "T2"), **props)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        sim.jacobian(("T1", "T2"), **props)
    stop.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - wall0) / steps
    work = atoms * pulses
    gpu_ms = start.elapsed_time(stop) / steps
    flops = work * TWO_POOL_FLOPS_PER_ORDER * 10
    achieved = flops / (gpu_ms * 1e-3) / 1e12
    return {"metric": "atom*pulses/s (two-pool signal+Jacobian)", "value": work / wall, "unit": "atom*pulses/s",
            "ms_per_step": wall * 1e3, "gpu_ms": gpu_ms, "dtype": "f32",
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS, "flops_per_atom_pulse": TWO_POOL_FLOPS_PER_ORDER * 10},
            "config": {"workload": "two-pool MRF, complex kernel", "atoms": atoms, "pulses": pulses, "states": 10,
                       "columns": ["T1", "T2"]}}


def semisolid_workload(steps: int = 5, warmup: int = 2) -> dict:
    """The semisolid (magnetization-transfer) pool and slice-profiled pulses
    on their own kernel, timed per call: an MT MRF dictionary (65,536 atoms x
    880 pulses, 10 states, signal + d/d(T1, T2, B1)); the same train with a
    windowed sinc worked out at 5 slice positions (8,192 voxels, signal +
    d/d(T1, T2)); and, at a design shape (1,024 atoms x 880 pulses, MT),
    loss.backward() to the flips, TRs and phases (adjoint) and one forward-mode
    tangent along them (dual tensors)."""
    import torch
    import torch.autograd.forward_ad as fwAD

    from torchsim_amd.sequence._transition import exact_slice_profile, sinc_pulse
    from torchsim_amd.simulators import MRFSimulator

    g = np.random.default_rng(0)
    device = torch.device("cuda", 0)
    train = torch.tensor(flips(), dtype=torch.float32)
    pulses = train.numel()

    def pool(atoms):
        return {"T1": torch.tensor(g.uniform(300, 2500, atoms), dtype=torch.float32, device=device),
                "T2": torch.tensor(g.uniform(30, 250, atoms), dtype=torch.float32, device=device),
                "B1": torch.tensor(g.uniform(0.8, 1.2, atoms), dtype=torch.float32, device=device),
                "bound_fraction": torch.tensor(g.uniform(0.05, 0.2, atoms), dtype=torch.float32, device=device),
                "bound_exchange": torch.tensor(g.uniform(5, 40, atoms), dtype=torch.float32, device=device),
                "bound_T1": torch.tensor(g.uniform(300, 1200, atoms), dtype=torch.float32, device=device)}

    def timed(call):
        for _ in range(warmup):
            call()
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        wall0 = time.perf_counter()
        start.record()
        for _ in range(steps):
            call()
        stop.record()
        torch.cuda.synchronize()
        return (time.perf_counter() - wall0) / steps, start.elapsed_time(stop) / steps

    atoms = 65536
    dictionary_maps = pool(atoms)
    mt = MRFSimulator(flip=train, TR=10.0, states=10)
    wall, gpu = timed(lambda: mt.jacobian(("T1", "T2", "B1"), **dictionary_maps))
    out = {"mt_dictionary": {"atoms": atoms, "pulses": pulses, "states": 10, "columns": ["T1", "T2", "B1"],
                             "ms_per_call": wall * 1e3, "gpu_ms": gpu, "atom_pulses_per_s": atoms * pulses / wall}}

    voxels, positions = 8192, 5
    plain = {k: dictionary_maps[k][:voxels] for k in ("T1", "T2")}
    profiled = MRFSimulator(flip=train, TR=10.0, states=10, pulse=sinc_pulse(2e-3, 4.0, raster_s=1e-5),
                            across_slice=exact_slice_profile(positions, bins=64))
    wall, gpu = timed(lambda: profiled.jacobian(("T1", "T2"), **plain))
    out["profiled_dictionary"] = {"voxels": voxels, "positions": positions, "pulses": pulses, "states": 10,
                                  "columns": ["T1", "T2"], "ms_per_call": wall * 1e3, "gpu_ms": gpu,
                                  "atom_pulses_per_s": voxels * positions * pulses / wall}

    design = 1024
    design_maps = pool(design)
    tr = torch.full((pulses,), 10.0)
    phases = torch.tensor(g.uniform(0, 360, pulses), dtype=torch.float32)
    weight = torch.tensor(g.normal(size=(design, pulses)), dtype=torch.float32, device=device)

    def gradient():
        leaves = [v.clone().requires_grad_(True) for v in (train, tr, phases)]
        sig = MRFSimulator(flip=leaves[0], TR=leaves[1], phases=leaves[2], states=10).simulate(**design_maps)
        (sig.real * weight).sum().backward()
        return leaves[0].grad

    directions = [torch.tensor(g.normal(size=pulses), dtype=torch.float32) for _ in range(3)]

    def tangent():
        with fwAD.dual_level():
            duals = [fwAD.make_dual(v, d) for v, d in zip((train, tr, phases), directions)]
            sig = MRFSimulator(flip=duals[0], TR=duals[1], phases=duals[2], states=10).simulate(**design_maps)
            return fwAD.unpack_dual(sig).tangent

    for name, call in (("mt_design_gradient", gradient), ("mt_design_tangent", tangent)):
        wall, gpu = timed(call)
        out[name] = {"atoms": design, "pulses": pulses, "states": 10, "ms_per_call": wall * 1e3, "gpu_ms": gpu}
    return {"metric": "semisolid and profiled kernels, ms per call", "dtype": "f32", "data": "synthetic",
            "config": {"workload": "MT MRF dictionary, profiled MRF dictionary, MT design gradient and tangent"},
            **out}


def subspace_workload(steps: int = 5, warmup: int = 2, rank: int = 8) -> dict:
    """The config-1 dictionary (1,274,184 atoms x 880 pulses, S=20, signal +
    d/d(T1,T2,B1)) projected onto a rank-8 temporal basis inside the kernel,
    beside the same projection done as simulate-then-GEMM (full outputs in
    HBM, then rocBLAS)."""
    import torch

    from torchsim_amd import Subspace
    from torchsim_amd.simulators import MRFSimulator

    device = torch.device("cuda", 0)
    T1, T2, B1 = dictionary(12)
    atoms = T1.size
    t1, t2, b1 = (torch.tensor(v, device=device) for v in (T1, T2, B1))
    seq = MRFSimulator(flip=torch.tensor(flips(), dtype=torch.float32), TR=10.0, states=STATES)
    subspace = Subspace.fit(seq.simulate(T1=t1[::97], T2=t2[::97], B1=b1[::97]).cpu(), rank)
    held = Subspace(basis=subspace.basis.to(device), singular_values=subspace.singular_values)

    def fused():
        return seq.project(held, NAMES, T1=t1, T2=t2, B1=b1)

    def unfused():
        signal, jacobian = seq.jacobian(NAMES, T1=t1, T2=t2, B1=b1)
        return held.project(signal), held.project(jacobian)

    def timed(step):
        out = None
        for _ in range(warmup):
            out = step()
        del out
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream(device)
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        wall0 = time.perf_counter()
        start.record(stream)
        for _ in range(steps):
            out = step()
            del out
        stop.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - wall0) / steps, start.elapsed_time(stop) / steps

    wall, gpu_ms = timed(fused)
    wall_unfused, gpu_unfused = timed(unfused)
    work = atoms * PULSES
    # the recursion's flops plus the projection's: per sample, each of the
    # 1 + P real streams times `rank` complex basis entries, 2 FMA each
    flops = FLOPS_PER_ATOM_PULSE + 4 * (1 + len(NAMES)) * rank
    achieved = flops * work / (gpu_ms * 1e-3) / 1e12
    return {"metric": "atom*pulses/s (signal+Jacobian projected to a rank-8 basis)", "value": work / wall,
            "unit": "atom*pulses/s", "ms_per_step": wall * 1e3, "gpu_ms": gpu_ms, "dtype": "f32",
            "unfused_ms_per_step": wall_unfused * 1e3, "unfused_gpu_ms": gpu_unfused,
            "config": {"workload": "MRF dictionary projected in the kernel", "atoms": atoms, "pulses": PULSES,
                       "states": STATES, "rank": rank, "columns": list(NAMES),
                       "outputs": "complex64 coefficients (A,8) + Jacobian coefficients (A,3,8)"},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS, "flops_per_atom_pulse": flops}}


def match_workload(steps: int = 3, warmup: int = 1, voxels: int = 65536) -> dict:
    """Dictionary matching: the MRF dictionary of the headline workload
    (1,274,184 atoms x 880 frames, complex64, real-subspace signal) adopted
    once, then `voxels` voxels (one 256 x 256 slice) matched against every
    atom per step; each voxel is an atom with a random complex receive gain
    plus noise at SNR 50. Reports voxel-atom scores per second and the MFMA
    rate the candidate kernel sustains (three fp16 products per FP32-accurate
    product)."""
    import torch

    from torchsim_amd.estimators import DictionaryMatcher
    from torchsim_amd.simulators import MRFSimulator

    device = torch.device("cuda", 0)
    T1, T2, B1 = dictionary(TILES)
    t1, t2, b1 = (torch.tensor(a, device=device) for a in (T1, T2, B1))
    sim = MRFSimulator(flip=torch.tensor(flips(), dtype=torch.float32), TR=10.0, states=STATES)
    atoms = sim.simulate(T1=t1, T2=t2, B1=b1)
    matcher = DictionaryMatcher(dictionary=atoms, parameters=torch.stack([t1, t2, b1], dim=-1))
    g = torch.Generator(device=device).manual_seed(0)
    pick = torch.randint(0, atoms.shape[0], (voxels,), device=device, generator=g)
    phase = torch.rand(voxels, device=device, generator=g) * 2 * np.pi
    gain = torch.polar(0.5 + 0.5 * torch.rand(voxels, device=device, generator=g), phase)
    sigma = float(atoms.abs().max()) / 50.0
    signals = gain[:, None] * atoms[pick] + sigma * torch.randn(voxels, atoms.shape[1], dtype=torch.complex64,
                                                               device=device, generator=g)
    del t1, t2, b1
    for _ in range(warmup):
        result = matcher.match(signals)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(device)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        result = matcher.match(signals)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - wall0) / steps
    gpu_s = start.elapsed_time(stop) * 1e-3 / steps
    pairs = voxels * atoms.shape[0]
    # the candidate pass, algorithmic: 2 rows per voxel x contrasts (one plane:
    # the MRF dictionary is real) x atoms x 3 fp16 products, 2 flops each
    # (padding of the operands to the tiles is not counted)
    mfma_flops = 3 * 2 * (2 * voxels) * atoms.shape[0] * atoms.shape[1]
    # the grid is tiled, so an atom's parameters repeat: compare those
    hit = float((result.parameters[:, 0, :] == matcher.parameter_values[pick]).all(-1).float().mean())
    return {"metric": "voxel*atom scores/s (dictionary matching)", "value": pairs / wall, "unit": "pairs/s",
            "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
            "dtype": "f32 (fp16x3 split products, FP64 refine)", "data": "synthetic",
            "config": {"workload": "MRF dictionary matching", "atoms": int(atoms.shape[0]),
                       "contrasts": int(atoms.shape[1]), "voxels": voxels, "top_k": 1},
            "roofline": {"bound": "mfma", "achieved": mfma_flops / gpu_s / 1e12, "peak": FP16_MFMA_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": mfma_flops / gpu_s / 1e12 / FP16_MFMA_PEAK_TFLOPS,
                         "traffic": None, "gpu_ms": gpu_s * 1e3},
            "parameters_recovered": hit}


def config0_workload(repeats: int = 3, warmup_s: float = 2.0) -> dict:
    """Plain ``mrf_sim`` at its CPU-runnable size: 1,000 atoms (T1 U[200,3000],
    T2 U[20,300] with T2 < T1, seed 0) x the 880-pulse train, TR 10 ms, S = 10,
    signal + d/d(T1, T2). Times the reference's own ``_epg_cpu`` kernel
    (built from source into oracle/_ref) the way its Jacobian runs it -- one
    simulate + simulate_jvp per property -- on one thread and on every host
    core (2 s warm-up, best and median of ``repeats``), and, where a GPU is
    visible, ``mrf_sim`` on it with the tissue already resident."""
    import statistics

    import torch

    out: dict = {"metric": "seconds per mrf_sim signal + Jacobian (T1, T2), 1,000 atoms x 880 pulses, S=10",
                 "unit": "s", "higher_is_better": False, "config": {"workload": "mrf_sim 1,000 x 880, S=10, d/d(T1,T2)"}}
    t1, t2 = random_tissue(1000, seed=0)
    train = flips()
    try:
        import oracle
        from torchsim_amd.simulators import MRFSimulator

        module = oracle.reference_module()
    except Exception:  # pragma: no cover - a broken build is reported as absent
        module = None
    if module is not None:
        events = oracle.events_of(MRFSimulator().packed(flip=torch.tensor(train, dtype=torch.float32), TR=10.0))
        tissue = {"t1": t1, "t2": t2}

        def once(threads):
            for name in ("t1", "t2"):
                oracle.reference_real(events, tissue, states=10, outputs=PULSES, features=0, threads=threads)
                oracle.reference_real(events, tissue, states=10, outputs=PULSES, features=0, threads=threads,
                                      tangent={name: np.ones(t1.size, dtype=np.float32)})

        cores = len(os.sched_getaffinity(0))
        for threads in sorted({1, cores}):
            begin = time.perf_counter()
            while time.perf_counter() - begin < warmup_s:
                once(threads)
            times = []
            for _ in range(repeats):
                start = time.perf_counter()
                once(threads)
                times.append(time.perf_counter() - start)
            out[f"reference_cpu_{threads}_threads"] = {"best_s": min(times), "median_s": statistics.median(times),
                                                        "atom_pulses_per_s": 1000 * PULSES / min(times)}
    if torch.cuda.is_available():
        from torchsim_amd import mrf_sim

        dev = torch.device("cuda", 0)
        T1, T2 = torch.tensor(t1, device=dev), torch.tensor(t2, device=dev)
        flip = torch.tensor(train, dtype=torch.float32)
        for _ in range(3):
            mrf_sim(flip, 10.0, T1, T2, diff=("T1", "T2"))
        torch.cuda.synchronize()
        times = []
        for _ in range(max(repeats, 10)):
            start = time.perf_counter()
            mrf_sim(flip, 10.0, T1, T2, diff=("T1", "T2"))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - start)
        out["gpu"] = {"best_s": min(times), "median_s": statistics.median(times),
                      "atom_pulses_per_s": 1000 * PULSES / min(times)}
    return out


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--tiles", type=int, default=TILES, help="copies of the 106,182-atom grid")
    parser.add_argument("--tile", default="auto",
                        help="state tile index, -1 for the library's cost model, or 'auto' (measured choice)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-torch-baseline", action="store_true")
    parser.add_argument("--workload", choices=("dictionary", "design", "fit", "serve", "match", "complex", "subspace", "semisolid",
                                                 "config0"), default="dictionary")
    parser.add_argument("--serve-seconds", type=float, default=60.0)
    parser.add_argument("--no-serving", action="store_true",
                        help="skip the short serving run the dictionary workload attaches (p50/p99 latency)")
    parser.add_argument(
        "--calibrate-copy",
        type=int,
        default=0,
        metavar="BYTES",
        help="first run one device-to-device copy of BYTES (a known-size kernel for counter scaling)",
    )
    args = parser.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
            "(atoms shard across GPUs with no exchange: run one process per GPU on its share)",
            file=sys.stderr,
        )
        return 2
    if args.workload == "serve":
        # many concurrent requests: more hardware queues than HIP's default 4
        # (read when the HIP runtime starts, so before torch or the library load)
        os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

    if args.workload == "config0":
        print(json.dumps(config0_workload()), flush=True)
        return 0

    import torch

    from torchsim_amd.simulators import MRFSimulator

    if args.workload == "design":
        print(json.dumps(design_workload(args.steps, args.warmup)), flush=True)
        return 0
    if args.workload == "fit":
        print(json.dumps(fit_workload()), flush=True)
        return 0
    if args.workload == "serve":
        print(json.dumps(serve_workload(duration_s=args.serve_seconds)), flush=True)
        return 0
    if args.workload == "complex":
        print(json.dumps(complex_workload(steps=args.steps, warmup=args.warmup)), flush=True)
        return 0
    if args.workload == "match":
        print(json.dumps(match_workload(steps=args.steps, warmup=args.warmup)), flush=True)
        return 0
    if args.workload == "subspace":
        print(json.dumps(subspace_workload(steps=args.steps, warmup=args.warmup)), flush=True)
        return 0
    if args.workload == "semisolid":
        print(json.dumps(semisolid_workload(steps=args.steps, warmup=args.warmup)), flush=True)
        return 0

    device = torch.device("cuda", 0)
    if args.calibrate_copy:
        # a kernel of known traffic for the HBM counters: reads and writes
        # calibrate_copy bytes once each (an elementwise multiply, so it runs
        # on the shader engines, where FETCH_SIZE / WRITE_SIZE count)
        source = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device=device)
        target = torch.empty_like(source)
        torch.mul(source, 2.0, out=target)
        torch.cuda.synchronize()
        del source, target
    tile = args.tile if args.tile == "auto" else int(args.tile)
    T1, T2, B1 = dictionary(args.tiles)
    atoms = T1.size
    t1 = torch.tensor(T1, device=device)
    t2 = torch.tensor(T2, device=device)
    b1 = torch.tensor(B1, device=device)
    seq = MRFSimulator(flip=torch.tensor(flips(), dtype=torch.float32), TR=10.0, states=STATES)

    def step():
        return seq.jacobian(NAMES, T1=t1, T2=t2, B1=b1, tile=tile)

    out = None
    for _ in range(args.warmup):
        out = step()
    del out
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(device)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        out = step()
        ends[k].record(stream)
        del out
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    kernel_ms = sum(s.elapsed_time(e) for s, e in zip(starts, ends)) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    work = atoms * PULSES
    value = work / (ms_per_step * 1e-3)
    achieved = FLOPS_PER_ATOM_PULSE * work / (kernel_ms * 1e-3) / 1e12
    record = {
        "metric": "atom*pulses/s (signal+Jacobian)",
        "value": value,
        "unit": "atom*pulses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "MRF dictionary signal + Jacobian d/d(T1,T2,B1)",
            "atoms": atoms,
            "pulses": PULSES,
            "states": STATES,
            "grid": "T1 100-5000 log x100, T2 10-2000 log x80 (T2<T1), B1 0.60:0.05:1.40, tiled x%d" % args.tiles,
            "outputs": "complex64 signal (A,880) + Jacobian (A,3,880)",
        },
        "roofline": {
            "bound": "valu",
            "bound_note": "FP32 vector ALU (an EPG recursion has no contraction for MFMA); peak = MI355X published FP32",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": measured_traffic(),
            "kernel_ms": kernel_ms,
            "flops_per_atom_pulse": FLOPS_PER_ATOM_PULSE,
            "write_GBs": BYTES_PER_ATOM_PULSE * work / (kernel_ms * 1e-3) / 1e9,
            "hbm_peak_GBs": HBM_PEAK_GBS,
        },
    }
    if not args.no_cpu_baseline:
        baseline = cpu_baseline()
        record["cpu_baseline"] = baseline
        record["speedup_vs_cpu_baseline"] = value / baseline["value"]
    if not args.no_torch_baseline:
        per_op = torch_per_op_baseline()
        record["torch_per_op_baseline"] = per_op
        record["speedup_vs_torch_per_op"] = value / per_op["value"]
    if not args.no_serving:
        record["serving"] = serving_summary()
    print(json.dumps(record), flush=True)
    return 0


def serving_summary(window_s: float = 10.0, limit_s: float = 240.0) -> dict:
    """The serving loop's p50/p99 latency over a short window of the Poisson
    mix (``--workload serve``), run as a child process after the timed
    region so that nothing it does can touch the headline figure."""
    import subprocess

    command = [sys.executable, os.path.abspath(__file__), "--workload", "serve", "--serve-seconds", str(window_s)]
    try:
        done = subprocess.run(command, capture_output=True, text=True, timeout=limit_s)
    except subprocess.TimeoutExpired:
        return {"error": f"no result within {limit_s:.0f} s"}
    lines = [line for line in done.stdout.splitlines() if line.startswith("{")]
    if done.returncode != 0 or not lines:
        return {"error": f"exit status {done.returncode}", "stderr": done.stderr[-400:]}
    served = json.loads(lines[-1])
    keep = ("p50_s", "p99_s", "by_kind", "value", "completed", "errors", "launches", "batched_jobs", "seconds")
    summary = {key: served.get(key) for key in keep}
    summary["req_per_s"] = summary.pop("value")
    summary["config"] = served.get("config")
    return summary


def measured_traffic():
    """HBM bytes per launch of the dictionary kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (each calibrated against a kernel
    of known traffic), or None when that summary is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "real_kernel_traffic.json")
    try:
        with open(path, encoding="utf-8") as handle:
            summary = json.load(handle)
        return float(summary["fetch_bytes"]) + float(summary["write_bytes"])
    except (OSError, ValueError, KeyError):
        return None


if __name__ == "__main__":
    sys.exit(main())
