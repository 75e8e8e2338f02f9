"""Off-resonance and a field spread on the samples, on the GPU: the kernels
run without B0 (the real-subspace kernel for a gradient echo, the complex
kernel for a quarter-turn spin echo) and the sample factor carries B0 and T2'
(reference sequence/_accelerators.py:459-497, 647-652, 736-737). Compared
with the FP64 oracle carrying B0 through every operator, times the spread's
exp(-1000 |tau| / T2') at the stream's known unrefocused times (TE for the
gradient echo; -2.5, 0, +3.5 ms about each spin echo) -- in value, in every
Jacobian column (T1, T2, B0, T2'), under loss.backward() and under
torch.func.jvp."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import packed as P
from tests import _cases

pytestmark = pytest.mark.gpu

TOL = 1e-5      # signal, relative to its largest sample (FP32)
GRAD_TOL = 2e-5  # derivatives, relative to their largest entry

ATOMS = 16
STATES = 48


def _case(name):
    if name == "gradient echo":
        return _cases.gradient_echo(np.linspace(10.0, 30.0, 24), 12.0, 3.0, states=STATES), np.full(24, 3e-3)
    sim = _cases.sampled_spin_echo(states=STATES)
    return sim, np.tile([-2.5e-3, 0.0, 3.5e-3], 4)


def _tissue(seed):
    rng = np.random.default_rng(seed)
    t1, t2 = _cases.random_tissue(ATOMS, seed=seed)
    return {"T1": t1, "T2": t2, "B0": rng.uniform(-40.0, 40.0, ATOMS).astype(np.float32),
            "T2prime": rng.uniform(10.0, 200.0, ATOMS).astype(np.float32)}


def _expected(sim, tau, maps, requires_grad=False):
    """FP64: the oracle with B0 in the states, times the spread's damping."""
    packed = sim.packed(**sim.bound)
    events = oracle.events_of(packed)
    values = {}
    for k, v in maps.items():
        value = torch.as_tensor(v, dtype=torch.float64)  # a tensor keeps its graph
        values[k] = value.clone().requires_grad_(True) if requires_grad else value
    signal = P.simulate(events, {"t1": values["T1"], "t2": values["T2"], "b0": values["B0"]}, states=STATES,
                        outputs=packed.output_count)
    damping = torch.exp(-(1000.0 / values["T2prime"])[:, None] * torch.as_tensor(np.abs(tau))[None, :])
    return signal * damping * complex(sim.signal_factor), values


def _rel(x, ref):
    return float(np.abs(x - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize("name", ["gradient echo", "spin echo"])
def test_signal_and_jacobian_against_b0_in_the_states(name):
    sim, tau = _case(name)
    maps = _tissue(5)
    gpu = {k: torch.tensor(v).cuda() for k, v in maps.items()}
    names = ("T1", "T2", "B0", "T2prime")
    signal, jacobian = sim.jacobian(names, **gpu)
    expect, values = _expected(sim, tau, maps)
    assert _rel(signal.cpu().numpy(), expect.numpy()) < TOL
    errors = {}
    for k, n in enumerate(names):
        column = torch.autograd.functional.jvp(
            lambda v: _expected(sim, tau, {**maps, n: v})[0], torch.tensor(maps[n], dtype=torch.float64),
            torch.ones(ATOMS, dtype=torch.float64))[1]
        errors[n] = _rel(jacobian[:, k].cpu().numpy(), column.numpy())
    print(name, {k: f"{v:.2e}" for k, v in errors.items()})
    assert max(errors.values()) < GRAD_TOL, errors
    assert torch.equal(sim.simulate(**gpu), signal)


@pytest.mark.parametrize("name", ["gradient echo", "spin echo"])
def test_loss_backward_reaches_b0_and_t2_prime(name):
    sim, tau = _case(name)
    maps = _tissue(6)
    rng = np.random.default_rng(7)
    weight = rng.normal(size=(ATOMS, len(tau))) + 1j * rng.normal(size=(ATOMS, len(tau)))
    gpu = {k: torch.tensor(v, device="cuda", requires_grad=True) for k, v in maps.items()}
    signal = sim.simulate(**gpu)
    (signal * torch.tensor(weight, dtype=torch.complex64, device="cuda").conj()).real.sum().backward()
    expect, values = _expected(sim, tau, maps, requires_grad=True)
    (expect * torch.tensor(weight).conj()).real.sum().backward()
    errors = {k: _rel(gpu[k].grad.cpu().numpy(), values[k].grad.numpy()) for k in maps}
    print(name, {k: f"{v:.2e}" for k, v in errors.items()})
    assert max(errors.values()) < GRAD_TOL, errors


def test_func_jvp_along_b0_and_t2_prime():
    sim, tau = _case("spin echo")
    maps = _tissue(8)
    rng = np.random.default_rng(9)
    direction = {"B0": rng.normal(size=ATOMS) * 5.0, "T2prime": rng.normal(size=ATOMS) * 3.0, "T1": rng.normal(size=ATOMS) * 50.0}
    keys = tuple(direction)
    primals = tuple(torch.tensor(maps[k]).cuda() for k in keys)
    tangents = tuple(torch.tensor(direction[k], dtype=torch.float32).cuda() for k in keys)
    others = {k: torch.tensor(v).cuda() for k, v in maps.items() if k not in keys}
    _, tangent = torch.func.jvp(lambda *v: sim.simulate(**others, **dict(zip(keys, v))), primals, tangents)
    _, expect = torch.autograd.functional.jvp(
        lambda *v: _expected(sim, tau, {**maps, **dict(zip(keys, v))})[0],
        tuple(torch.tensor(maps[k], dtype=torch.float64) for k in keys),
        tuple(torch.tensor(direction[k], dtype=torch.float64) for k in keys))
    error = _rel(tangent.cpu().numpy(), expect.numpy())
    print(f"jvp along B0, T2', T1: {error:.2e}")
    assert error < GRAD_TOL
