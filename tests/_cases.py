"""Workloads shared by the tests: the MRF trains and tissue draws of the
benchmark configurations, at sizes the FP64 oracle finishes in seconds."""

from __future__ import annotations

import numpy as np
import torch


def mrf_flips() -> np.ndarray:
    """880 pulses: 5->60 deg over 300, 60->2 over 300, then 280 at 2 deg."""
    return np.concatenate([np.linspace(5, 60, 300), np.linspace(60, 2, 300), np.full(280, 2.0)])


def benchmark_flips(length: int) -> np.ndarray:
    """The reference benchmark's train (benchmarks/_common.py:25-35)."""
    up = np.linspace(5.0, 60.0, round(0.375 * length))
    down = np.linspace(60.0, 2.0, round(0.375 * length))
    tail = np.full(length - up.size - down.size, 2.0)
    return np.concatenate((up, down, tail))


def benchmark_grid(atoms: int) -> tuple[np.ndarray, np.ndarray]:
    """Paired geomspace T1/T2 of the reference benchmark (benchmarks/_common.py:38-45)."""
    return np.geomspace(200.0, 3000.0, atoms), np.geomspace(10.0, 300.0, atoms)


def random_tissue(atoms: int, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """T1 ~ U[200, 3000] ms, T2 ~ U[20, 300] ms redrawn until T2 < T1."""
    generator = torch.Generator().manual_seed(seed)
    t1 = torch.empty(atoms, dtype=torch.float64).uniform_(200.0, 3000.0, generator=generator)
    t2 = torch.empty(atoms, dtype=torch.float64).uniform_(20.0, 300.0, generator=generator)
    bad = t2 >= t1
    while bool(bad.any()):
        t2[bad] = torch.empty(int(bad.sum()), dtype=torch.float64).uniform_(20.0, 300.0, generator=generator)
        bad = t2 >= t1
    return t1.numpy().astype(np.float32), t2.numpy().astype(np.float32)


def dictionary_grid() -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """T1 log 100-5000 (100) x T2 log 10-2000 (80), T2 < T1, x B1 0.60:0.05:1.40."""
    t1 = np.geomspace(100.0, 5000.0, 100)
    t2 = np.geomspace(10.0, 2000.0, 80)
    b1 = np.round(np.arange(0.60, 1.40 + 1e-9, 0.05), 2)
    T1, T2 = np.meshgrid(t1, t2, indexing="ij")
    keep = T2 < T1
    pairs_t1, pairs_t2 = T1[keep], T2[keep]
    T1b = np.repeat(pairs_t1[:, None], b1.size, 1).reshape(-1)
    T2b = np.repeat(pairs_t2[:, None], b1.size, 1).reshape(-1)
    B1b = np.tile(b1, pairs_t1.size)
    return T1b.astype(np.float32), T2b.astype(np.float32), B1b.astype(np.float32)


def rel_err(x: np.ndarray, ref: np.ndarray) -> float:
    """max|x - ref| / max|ref|, the reference's convention (test_cuda_parity.py:131-133)."""
    return float(np.abs(x - ref).max() / max(np.abs(ref).max(), 1e-300))


def phantom(shape=(256, 256, 8), *, classes=6, seed=0, device="cpu"):
    """Piecewise-smooth T1/T2/M0 maps: ``classes`` random ellipsoidal tissue
    classes (the first fills the field of view) with T1 300-2500 ms, T2
    30-250 ms, M0 0.5-1.0, each map times (1 + 5% Gaussian-smoothed noise)."""
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
        inside = (((grid - center) / radii) ** 2).sum(-1) <= 1.0
        label[inside] = k
    maps = []
    for values in (t1c, t2c, m0c):
        base = values[label]
        noise = torch.randn(shape, generator=generator, dtype=torch.float64)
        kernel = torch.exp(-0.5 * (torch.arange(-6, 7, dtype=torch.float64) / 2.0) ** 2)
        kernel = kernel / kernel.sum()
        smooth = noise[None, None]
        for axis in range(3):
            view = [1, 1, 1]
            view[axis] = kernel.numel()
            pad = [0, 0, 0, 0, 0, 0]
            pad[2 * (2 - axis)] = pad[2 * (2 - axis) + 1] = 6
            smooth = torch.nn.functional.conv3d(torch.nn.functional.pad(smooth, pad, mode="replicate"),
                                                kernel.reshape(1, 1, *view))
        smooth = smooth[0, 0]
        smooth = smooth / smooth.abs().max()
        maps.append((base * (1.0 + 0.05 * smooth)).to(torch.float32).to(device))
    return tuple(maps)


def gradient_echo(flips_deg, tr_ms, te_ms, *, spoiled=True, states=48):
    """A gradient-echo train as a simulator: each pulse, its sample TE later,
    then ideal spoiling (or one unbalanced gradient), then the rest of TR."""
    from torchsim_amd.model._simulator import Simulator, SpinPhysics
    from torchsim_amd.sequence._events import Delay, Dephase, Excitation, Readout, Spoil, module

    class GradientEcho(Simulator):
        model = SpinPhysics(properties={"T1": "t1_ms", "T2": "t2_ms", "M0": "m0", "B1": "b1"})
        signal_factor = 1.0

        def layout(self, *, flip, TR, TE, spoiled=True):
            angles = np.deg2rad(np.atleast_1d(np.asarray(flip, dtype=np.float64)))
            trs = np.broadcast_to(np.asarray(TR, dtype=np.float64), angles.shape)
            parts = []
            for angle, tr in zip(angles, trs):
                after = Spoil() if spoiled else Dephase()
                parts.append(module(Excitation(float(angle)), Delay(TE * 1e-3), Readout(0.0), after,
                                    Delay((tr - TE) * 1e-3), duration_s=tr * 1e-3))
            return parts

    GradientEcho.states = states
    return GradientEcho(flip=np.asarray(flips_deg, dtype=np.float64), TR=tr_ms, TE=float(te_ms),
                        spoiled=bool(spoiled))


def sampled_spin_echo(echoes=4, spacing_ms=8.0, offsets_ms=(-2.5, 0.0, 3.5), states=48):
    """A refocused train (90 deg at a quarter turn, crushed 180s) sampled off
    the echo centres as well as on them."""
    from torchsim_amd.model._simulator import Simulator, SpinPhysics
    from torchsim_amd.sequence._events import Excitation, Readout, Refocusing

    class SampledSpinEcho(Simulator):
        model = SpinPhysics(properties={"T1": "t1_ms", "T2": "t2_ms", "M0": "m0"})
        signal_factor = 1.0

        def layout(self, *, echoes, ESP, offsets):
            parts = [(0.0, Excitation(np.pi / 2, np.pi / 2))]
            for k in range(echoes):
                echo = (k + 1) * ESP * 1e-3
                parts.append((echo - 0.5 * ESP * 1e-3, Refocusing(np.pi, 0.0)))
                parts += [(echo + offset * 1e-3, Readout(0.0)) for offset in offsets]
            return parts

    SampledSpinEcho.states = states
    return SampledSpinEcho(echoes=int(echoes), ESP=float(spacing_ms), offsets=tuple(float(o) for o in offsets_ms))
