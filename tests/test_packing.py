"""The differentiable MRF packing used for sequence design: same buffers as
the generic event-by-event packing, bit for bit, and the exact derivatives of
the intervals and flips with respect to the TRs and flip angles."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from torchsim_amd.model._simulator import Simulator
from torchsim_amd.simulators import MRFSimulator

FIELDS = ("duration", "flip", "phase", "kind", "action", "output_index")


@pytest.mark.parametrize("pulses, ti", [(7, 0.0), (300, 0.0), (64, 20.0)])
def test_vectorised_packing_matches_generic(pulses, ti):
    rng = np.random.default_rng(pulses)
    flips, trs = rng.uniform(2.0, 70.0, pulses), rng.uniform(8.0, 15.0, pulses)
    seq = MRFSimulator(states=10)
    flip_a = torch.tensor(flips, dtype=torch.float32, requires_grad=True)
    tr_a = torch.tensor(trs, dtype=torch.float32, requires_grad=True)
    fast = seq._packed(flip=flip_a, TR=tr_a, TI=ti)
    flip_b = torch.tensor(flips, dtype=torch.float32, requires_grad=True)
    tr_b = torch.tensor(trs, dtype=torch.float32, requires_grad=True)
    generic = Simulator._packed(seq, flip=flip_b, TR=tr_b, TI=ti)
    for field in FIELDS:
        assert torch.equal(getattr(fast, field).detach(), getattr(generic, field).detach()), field
    assert fast.output_count == generic.output_count
    w = torch.randn(fast.duration.numel())
    v = torch.randn(fast.flip.numel())
    ((fast.duration * w).sum() + (fast.flip * v).sum()).backward()
    ((generic.duration * w).sum() + (generic.flip * v).sum()).backward()
    assert torch.allclose(flip_a.grad, flip_b.grad, rtol=1e-6, atol=0.0)
    # the generic graph rounds through the float32 clock; the vectorised one is exact
    assert float((tr_a.grad - tr_b.grad).abs().max()) <= 1e-5 * float(tr_b.grad.abs().max())
    turning = ((fast.kind == 1) & ((fast.action & 4) == 0)).nonzero().reshape(-1)
    # interval of pulse k+1 is TR_k (ms -> s); the last TR is never played
    assert torch.allclose(tr_a.grad[:-1], 1e-3 * w[turning[1:]], rtol=1e-6, atol=0.0)
    assert float(tr_a.grad[-1]) == 0.0
    assert torch.allclose(flip_a.grad, torch.deg2rad(torch.ones(())) * v[turning], rtol=1e-6, atol=0.0)


def test_plain_protocols_still_take_the_generic_packing():
    seq = MRFSimulator(states=10)
    flips = torch.linspace(5.0, 60.0, 20)
    assert torch.equal(seq.packed(flip=flips, TR=10.0).duration, Simulator._packed(seq, flip=flips, TR=10.0).duration)


@pytest.mark.parametrize("simulator, protocol", [
    ("MRFSimulator", dict(flip=np.r_[np.linspace(5, 60, 300), np.linspace(60, 2, 300), np.full(280, 2.0)], TR=10.0)),
    ("MRFSimulator", dict(flip=np.linspace(3, 70, 200), TR=np.linspace(8, 15, 200), TI=20.0)),
    ("FSESimulator", dict(flip=np.linspace(180, 120, 96), ESP=5.0, TR=3000.0)),
    ("TrueFISPSimulator", dict(flip=35.0, TR=4.5, nshots=400)),
    ("TrueFISPSimulator", dict(flip=50.0, TR=4.5, nshots=300, inversion="ideal", spoiler=5.0)),
])
def test_host_scalar_layouts_pack_the_same_bits(simulator, protocol, monkeypatch):
    """Layouts walk NumPy scalars when no graph runs through the protocol;
    the buffers are those that 0-dim tensors give, bit for bit."""
    import torchsim_amd.simulators as sims
    from torchsim_amd.simulators import fse, mrf, truefisp

    sim = getattr(sims, simulator)()
    args = {k: (torch.tensor(v, dtype=torch.float32) if isinstance(v, np.ndarray) else v) for k, v in protocol.items()}
    fast = Simulator._packed(sim, **args)
    for module in (mrf, fse, truefisp):
        monkeypatch.setattr(module, "host_values", lambda *arrays: arrays)
    tensors = Simulator._packed(sim, **args)
    for field in FIELDS:
        assert torch.equal(getattr(fast, field), getattr(tensors, field)), field
    assert fast.output_count == tensors.output_count


@pytest.mark.parametrize("pulses, tr, ti, phases", [
    (1, 10.0, 0.0, 0.0), (2, 10.0, 0.0, 0.0), (3, 12.5, 5.0, 0.0), (17, "ramp", 0.0, 0.3),
    (880, 10.0, 0.0, 0.0), (750, "ramp", 20.0, 0.0),
])
def test_tiled_packing_of_plain_protocols_matches_generic(pulses, tr, ti, phases):
    """Protocols without a graph pack from the cached two-pulse template,
    tiled: every buffer equals the event-by-event packing bit for bit."""
    rng = np.random.default_rng(pulses)
    flips = torch.tensor(rng.uniform(2.0, 70.0, pulses), dtype=torch.float32)
    trs = torch.tensor(np.linspace(8.0, 15.0, pulses), dtype=torch.float32) if tr == "ramp" else tr
    seq = MRFSimulator(states=10)
    fast = seq._packed(flip=flips, TR=trs, TI=ti, phases=phases)
    generic = Simulator._packed(seq, flip=flips, TR=trs, TI=ti, phases=phases)
    for field in FIELDS + ("shim_index", "saturation", "rf_frequency_hz"):
        assert torch.equal(getattr(fast, field), getattr(generic, field)), field
    assert fast.output_count == generic.output_count == pulses


@pytest.mark.parametrize("pulses, ti, shaped", [(1, 0.0, False), (5, 0.0, False), (880, 0.0, False), (64, 20.0, False),
                                               (40, 0.0, True)])
def test_per_pulse_phases_pack_from_the_template(pulses, ti, shaped):
    """A phase-cycled protocol (per-pulse float32 phases, plain or carrying
    a graph, with a shaped pulse too) packs from the template: every buffer
    equals the event-by-event packing bit for bit, and the phases' and
    flips' gradients are the generic graph's."""
    from torchsim_amd.sequence._transition import sinc_pulse

    rng = np.random.default_rng(pulses + 1)
    flips, trs = rng.uniform(2.0, 70.0, pulses), rng.uniform(8.0, 15.0, pulses)
    phases = rng.uniform(-360.0, 360.0, pulses)
    seq = MRFSimulator(states=10)
    extra = {"pulse": sinc_pulse(2e-3, 4.0, raster_s=1e-5)} if shaped else {}
    plain = {"flip": torch.tensor(flips, dtype=torch.float32), "TR": torch.tensor(trs, dtype=torch.float32),
             "phases": torch.tensor(phases, dtype=torch.float32), "TI": ti, **extra}
    fast = seq._packed(**plain)
    generic = Simulator._packed(seq, **plain)
    for field in FIELDS + ("shim_index", "saturation", "rf_frequency_hz"):
        assert torch.equal(getattr(fast, field), getattr(generic, field)), field
    assert seq._pulse_block({k: v for k, v in plain.items() if k not in ("flip", "TR")}) is not None

    def leaves():
        return [torch.tensor(v, dtype=torch.float32, requires_grad=True) for v in (flips, trs, phases)]

    a, b = leaves(), leaves()
    fast = seq._packed(flip=a[0], TR=a[1], phases=a[2], TI=ti, **extra)
    generic = Simulator._packed(seq, flip=b[0], TR=b[1], phases=b[2], TI=ti, **extra)
    for field in FIELDS:
        assert torch.equal(getattr(fast, field).detach(), getattr(generic, field).detach()), field
    w, v, u = (torch.randn(fast.duration.numel()) for _ in range(3))
    ((fast.duration * w).sum() + (fast.flip * v).sum() + (fast.phase * u).sum()).backward()
    ((generic.duration * w).sum() + (generic.flip * v).sum() + (generic.phase * u).sum()).backward()
    assert torch.allclose(a[0].grad, b[0].grad, rtol=1e-6, atol=0.0)
    assert torch.allclose(a[2].grad, b[2].grad, rtol=1e-6, atol=0.0)
    tr_generic = torch.zeros(pulses) if b[1].grad is None else b[1].grad  # one pulse plays no TR
    assert float((a[1].grad - tr_generic).abs().max()) <= 1e-5 * float(tr_generic.abs().max())
    # a pulse's phase reaches its excitation and its sample: d/dphase = (u_rf + u_adc) pi/180
    assert float(a[2].grad.abs().max()) > 0.0


def test_tiled_packing_is_fast():
    import time

    seq = MRFSimulator(states=10)
    flips = torch.linspace(5.0, 60.0, 1000)
    seq._packed(flip=flips, TR=10.0)  # the template, once
    start = time.perf_counter()
    for k in range(20):
        seq._packed(flip=flips + 0.01 * k, TR=10.0)
    assert (time.perf_counter() - start) / 20 < 0.01
