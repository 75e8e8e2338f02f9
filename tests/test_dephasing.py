"""Off-resonance and a field spread on the samples (no GPU): the unrefocused
time and the uniform-winding verdict read off the packed stream (reference
sequence/_accelerators.py:820-857, 950-1165; its tests
tests/sequence/test_unrefocused_time.py and test_static_spread.py), checked
against the FP64 oracle carrying B0 through every operator and against an
isochromat ensemble over a Cauchy spread of frequencies."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import packed as P
from tests import _cases

B0_HZ = 53.0  # not a whole number of turns per TR or echo spacing
T2_PRIME_MS = 25.0
LENGTH = 24


def _streams():
    from torchsim_amd.simulators import FSESimulator, MRFSimulator, TrueFISPSimulator

    flips = torch.full((LENGTH,), 40.0)
    ramp = torch.linspace(11.0, 15.0, LENGTH)
    uniform = {
        "spoiled, TE 3 ms": _cases.gradient_echo(np.full(LENGTH, 20.0), 12.0, 3.0),
        "spoiled, TE 6 ms": _cases.gradient_echo(np.full(LENGTH, 20.0), 12.0, 6.0),
        "spoiled, varying TR": _cases.gradient_echo(np.full(LENGTH, 20.0), np.linspace(11.0, 15.0, LENGTH), 3.0),
        "unbalanced gradient echo, TE 3 ms": _cases.gradient_echo(np.full(LENGTH, 20.0), 12.0, 3.0, spoiled=False),
        "spin echo sampled off the echoes": _cases.sampled_spin_echo(),
        "refocused, 180 deg": FSESimulator(flip=torch.full((LENGTH,), 180.0), ESP=5.0),
        "refocused, 60 deg": FSESimulator(flip=torch.full((LENGTH,), 60.0), ESP=5.0),
        "unbalanced, one TR": MRFSimulator(flip=flips, TR=10.0),
        "unbalanced, inversion prepared": MRFSimulator(flip=flips, TR=10.0, TI=20.0),
    }
    refused = {
        "unbalanced, ramped TR": MRFSimulator(flip=flips, TR=ramp),
        "unbalanced, random TR": MRFSimulator(
            flip=flips, TR=8.0 + 8.0 * torch.rand(LENGTH, generator=torch.Generator().manual_seed(0))),
        "unbalanced gradient echo, varying TR": _cases.gradient_echo(np.full(LENGTH, 20.0), np.linspace(11.0, 15.0, LENGTH),
                                                                     3.0, spoiled=False),
        "balanced": TrueFISPSimulator(flip=40.0, TR=5.0, nshots=LENGTH),
    }
    return uniform, refused


UNIFORM, REFUSED = _streams()


def _packed(sim):
    return sim.packed(**sim.bound)


def _oracle(packed, b0_hz, atoms=1, states=48):
    events = oracle.events_of(packed)
    tissue = {"t1": np.full(atoms, 1000.0), "t2": np.full(atoms, 80.0), "b0": b0_hz}
    return P.simulate(events, tissue, states=states, outputs=packed.output_count).numpy()


def test_unrefocused_times_of_known_streams():
    """A gradient echo carries its TE to every sample; a spin echo carries
    nothing to its echoes and the signed offset either side; a FISP read at
    once after its pulse carries nothing (float32 timestamps hold a
    millisecond to about a nanosecond)."""
    from torchsim_amd.sequence._dephasing import unrefocused_s

    for te, name in ((3e-3, "spoiled, TE 3 ms"), (6e-3, "spoiled, TE 6 ms"), (3e-3, "spoiled, varying TR"),
                     (3e-3, "unbalanced gradient echo, TE 3 ms")):
        tau = unrefocused_s(_packed(UNIFORM[name])).double().numpy()
        assert tau.shape == (LENGTH,) and np.abs(tau - te).max() < 1e-5 * te, name
    for name in ("refocused, 180 deg", "refocused, 60 deg", "unbalanced, one TR", "unbalanced, inversion prepared"):
        assert np.abs(unrefocused_s(_packed(UNIFORM[name])).numpy()).max() < 1e-5 * 5e-3, name
    tau = unrefocused_s(_packed(UNIFORM["spin echo sampled off the echoes"])).double().numpy().reshape(4, 3)
    assert np.abs(tau - np.array([-2.5e-3, 0.0, 3.5e-3])).max() < 1e-8


@pytest.mark.parametrize("name", list(UNIFORM))
def test_the_turn_on_the_samples_is_what_the_states_compute(name):
    """B0 through every operator of the FP64 oracle equals the B0-free
    signal turned by exp(-2 pi i B0 tau) (float32 timestamps leave ~1e-6)."""
    from torchsim_amd.sequence._dephasing import unrefocused_s, winding

    packed = _packed(UNIFORM[name])
    assert winding(packed).analytic
    tau = unrefocused_s(packed).double().numpy()
    still = _oracle(packed, 0.0)[0]
    turned = _oracle(packed, B0_HZ)[0]
    predicted = still * np.exp(-2j * np.pi * B0_HZ * tau)
    assert np.abs(turned - predicted).max() < 1e-5 * np.abs(still).max()


@pytest.mark.parametrize("name", list(REFUSED))
def test_a_refused_stream_is_one_the_turn_would_get_wrong(name):
    """The verdict is not caution: where it refuses, the one-number turn
    misses what the states compute by more than a percent."""
    from torchsim_amd.sequence._dephasing import unrefocused_s, winding

    packed = _packed(REFUSED[name])
    assert not winding(packed).analytic
    tau = unrefocused_s(packed).double().numpy()
    still = _oracle(packed, 0.0)[0]
    turned = _oracle(packed, B0_HZ)[0]
    assert np.abs(turned - still * np.exp(-2j * np.pi * B0_HZ * tau)).max() > 1e-2 * np.abs(still).max()


@pytest.mark.parametrize("plant", [None, "double_rate"])
@pytest.mark.parametrize("name", ["spoiled, TE 3 ms", "spin echo sampled off the echoes"])
def test_a_lorentzian_spread_is_an_ensemble_of_frequencies(name, plant):
    """A Cauchy spread of half-width R2' / (2 pi) Hz, played through the
    oracle's states as 20,001 isochromats at its quantiles, averages to the
    spread-free signal times exp(-R2' |tau|), R2' = 1000 / T2' (quadrature
    to ~5e-4); a rate off by two (planted) misses by ~0.1."""
    from torchsim_amd.sequence._dephasing import Dephasing, unrefocused_s

    packed = _packed(UNIFORM[name])
    count = 20001
    quantiles = (np.arange(count) + 0.5) / count
    half_width_hz = (1000.0 / T2_PRIME_MS) / (2 * np.pi)
    ensemble = _oracle(packed, half_width_hz * np.tan(np.pi * (quantiles - 0.5)), atoms=count).mean(0)
    still = _oracle(packed, 0.0)[0]
    t2_prime = T2_PRIME_MS / (2.0 if plant else 1.0)
    factor = Dephasing(b0_hz=None, t2_prime_ms=torch.tensor([t2_prime]), tau_s=unrefocused_s(packed)).factor()
    error = np.abs(ensemble - still * factor.numpy()[0]).max() / np.abs(still).max()
    if plant is None:
        assert error < 2e-3
    else:
        assert error > 2e-2


def test_the_factor_turns_and_damps_together():
    """exp(-R2'|tau|) exp(-2 pi i B0 tau) per atom and sample, and the two
    closed-form Jacobian columns (-2 pi i tau s; |tau| 1000 / T2'^2 s)
    against central differences of the factor in FP64."""
    from torchsim_amd.sequence._dephasing import Dephasing

    tau = torch.tensor([-2.5e-3, 0.0, 3.5e-3, 1e-2], dtype=torch.float64)
    b0 = torch.tensor([0.0, 12.5, -40.0], dtype=torch.float64)
    t2p = torch.tensor([25.0, 80.0, 3.0], dtype=torch.float64)
    made = Dephasing(b0_hz=b0, t2_prime_ms=t2p, tau_s=tau).factor()
    expect = torch.exp(-(1000.0 / t2p)[:, None] * tau.abs()) * torch.exp(-2j * np.pi * b0[:, None] * tau)
    assert torch.allclose(made, expect, rtol=1e-12, atol=0)
    for name in ("b0_hz", "t2_prime_ms"):
        step = 1e-4 if name == "b0_hz" else 1e-4 * t2p

        def at(sign):
            moved = {"b0_hz": b0, "t2_prime_ms": t2p}
            moved[name] = moved[name] + sign * step
            return Dephasing(b0_hz=moved["b0_hz"], t2_prime_ms=moved["t2_prime_ms"], tau_s=tau).factor()

        numeric = (at(1) - at(-1)) / (2 * (step if name == "b0_hz" else step[:, None]))
        column = Dephasing(b0_hz=b0, t2_prime_ms=t2p, tau_s=tau).column(made, name)
        assert torch.allclose(column, numeric, rtol=1e-6, atol=1e-9), name


def test_the_simulator_holds_b0_out_where_the_winding_admits_it():
    """Uniform winding: B0 leaves the kernels' tissue (the run keeps the
    real-subspace kernel) and reaches the factor; a ramped TR keeps B0 in
    the states; a spread on a stream that cannot carry it is refused."""
    from torchsim_amd import _native
    from torchsim_amd.sequence import _engine

    maps = {"T1": torch.tensor([1000.0, 800.0]), "T2": torch.tensor([80.0, 60.0]), "B0": torch.tensor([10.0, -5.0])}
    sim = UNIFORM["unbalanced, one TR"]
    held, sequence = sim._split({**sim.bound, **maps, "T2prime": 30.0})
    _, packed, _, tissue, _, _, _, _, dephasing = sim._prepared(held, sequence, (), device="cpu")
    assert dephasing is not None and tissue.buffers[_native.TISSUE_INDEX["b0_hz"]] is None
    assert torch.equal(dephasing.b0_hz, maps["B0"]) and torch.equal(dephasing.t2_prime_ms, torch.full((2,), 30.0))
    assert _engine.real_subspace(packed, tissue)[0]

    ramped = REFUSED["unbalanced, ramped TR"]
    held, sequence = ramped._split({**ramped.bound, **maps})
    _, _, _, tissue, _, _, _, _, dephasing = ramped._prepared(held, sequence, (), device="cpu")
    assert dephasing is None and torch.equal(tissue.buffers[_native.TISSUE_INDEX["b0_hz"]], maps["B0"])
    held, sequence = ramped._split({**ramped.bound, **maps, "T2prime": 30.0})
    with pytest.raises(ValueError, match="unrefocused"):
        ramped._prepared(held, sequence, (), device="cpu")
    balanced = REFUSED["balanced"]
    held, sequence = balanced._split({**balanced.bound, "T1": 1000.0, "T2": 80.0, "T2prime": 30.0})
    with pytest.raises(ValueError, match="unrefocused"):
        balanced._prepared(held, sequence, (), device="cpu")
    # an infinite T2' (the identity) with no derivative asks for nothing
    held, sequence = ramped._split({**ramped.bound, "T1": 1000.0, "T2": 80.0, "T2prime": float("inf")})
    assert ramped._prepared(held, sequence, (), device="cpu")[-1] is None
