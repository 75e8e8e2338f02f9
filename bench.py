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


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=10)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--tiles", type=int, default=TILES, help="copies of the 106,182-atom grid")
    parser.add_argument("--tile", type=int, default=-1, help="state tile index (-1: library choice)")
    parser.add_argument("--no-cpu-baseline", action="store_true")
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
            "(atoms shard across GPUs with no exchange; see DESIGN.md)",
            file=sys.stderr,
        )
        return 2

    import torch

    from torchsim_amd.simulators import MRFSimulator

    device = torch.device("cuda", 0)
    if args.calibrate_copy:
        source = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device=device)
        target = torch.empty_like(source)
        target.copy_(source)
        torch.cuda.synchronize()
        del source, target
    T1, T2, B1 = dictionary(args.tiles)
    atoms = T1.size
    t1 = torch.tensor(T1, device=device)
    t2 = torch.tensor(T2, device=device)
    b1 = torch.tensor(B1, device=device)
    seq = MRFSimulator(flip=torch.tensor(flips(), dtype=torch.float32), TR=10.0, states=STATES)

    def step():
        return seq.jacobian(NAMES, T1=t1, T2=t2, B1=b1, tile=args.tile)

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
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None,
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
    print(json.dumps(record), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
