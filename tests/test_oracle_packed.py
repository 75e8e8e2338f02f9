"""Pins the FP64 complex oracle (oracle/packed.py)."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg
import torch

import oracle
from oracle import isochromat
from oracle import packed as P
from tests import _cases


def _mrf_events(flip, tr, **kw):
    from torchsim_amd.simulators import MRFSimulator

    return oracle.events_of(MRFSimulator().packed(flip=flip, TR=tr, **kw))


def _phased(events, seed=0):
    rng = np.random.default_rng(seed)
    out = dict(events)
    n = len(events["kind"])
    out["phase"] = np.where(events["kind"] == 1, rng.uniform(-3, 3, n), np.where(events["kind"] == 2, 0.4, 0.0))
    return out


def test_matches_c_oracle_in_the_real_subspace():
    events = _mrf_events(torch.tensor(_cases.benchmark_flips(80)), 9.0)
    t1, t2 = _cases.random_tissue(5, seed=1)
    b1 = np.linspace(0.8, 1.2, 5)
    ref, tan = oracle.real(events, t1=t1, t2=t2, b1=b1, states=12, outputs=80,
                           tangents=[{"t1": 1.0}, {"t2": 1.0}, {"b1": 1.0}])
    sig, jac = P.jacobian(events, {"t1": t1, "t2": t2, "b1": b1}, ("t1", "t2", "b1"), states=12, outputs=80)
    assert np.abs(sig.real.numpy()).max() < 1e-15
    assert np.abs(sig.imag.numpy() - ref).max() < 1e-13
    assert np.abs(jac.imag.numpy() - tan).max() < 1e-12


@pytest.mark.parametrize("b0", [0.0, 7.5])
def test_isochromat_brute_force_with_phases_and_off_resonance(b0):
    events = _phased(_mrf_events(torch.tensor(np.linspace(10, 70, 25)), torch.tensor(np.linspace(8, 14, 25))))
    sig = P.simulate(events, {"t1": 800.0, "t2": 60.0, "b1": 0.9, "b0": b0}, states=64, outputs=25)
    iso = isochromat.simulate(events, t1=800.0, t2=60.0, b1=0.9, b0=b0, isochromats=256)
    assert np.abs(sig.numpy()[0] - iso).max() < 1e-12


def test_exchange_with_zero_rate_is_two_independent_pools():
    events = _phased(_mrf_events(torch.tensor(_cases.benchmark_flips(60)), 10.0))
    f = 0.2
    two = P.simulate(events, {"t1": 900.0, "t2": 70.0, "fraction": f, "exchange": 0.0, "t1b": 400.0,
                              "t2b": 15.0, "shift": 30.0}, states=10, outputs=60)
    a = P.simulate(events, {"t1": 900.0, "t2": 70.0}, states=10, outputs=60)
    b = P.simulate(events, {"t1": 400.0, "t2": 15.0, "b0": 30.0}, states=10, outputs=60)
    assert np.abs(two.numpy() - ((1 - f) * a + f * b).numpy()).max() < 1e-10


def test_exchange_operators_against_scipy_expm():
    kab, kba, r1a, r1b = 20 * 0.2, 20 * 0.8, 1000 / 900, 1000 / 400
    g = np.array([[-kab - r1a, kba], [kab, -kba - r1b]])
    dt = 0.01
    ours = torch.linalg.matrix_exp(torch.tensor(g * dt)).numpy()
    assert np.abs(ours - scipy.linalg.expm(g * dt)).max() < 1e-15
    # the oracle's recovery is the affine solution of dz/dt = G z + s
    s = np.array([0.8 * r1a, 0.2 * r1b])
    z0 = np.array([-0.3, 0.1])
    from scipy.integrate import solve_ivp

    sol = solve_ivp(lambda t, z: g @ z + s, (0, dt), z0, rtol=1e-12, atol=1e-14)
    affine = scipy.linalg.expm(g * dt) @ z0 + (scipy.linalg.expm(g * dt) - np.eye(2)) @ np.linalg.solve(g, s)
    assert np.abs(sol.y[:, -1] - affine).max() < 1e-10


def test_exchange_conserves_total_without_relaxation():
    """With relaxation switched off, exchange between two equal-frequency pools
    leaves the summed signal identical to one pool."""
    events = _phased(_mrf_events(torch.tensor(_cases.benchmark_flips(40)), 10.0))
    big = 1e12
    two = P.simulate(events, {"t1": big, "t2": big, "fraction": 0.3, "exchange": 15.0, "t1b": big,
                              "t2b": big, "shift": 0.0}, states=10, outputs=40)
    one = P.simulate(events, {"t1": big, "t2": big}, states=10, outputs=40)
    assert np.abs(two.numpy() - one.numpy()).max() < 1e-9


def test_diffusion_spin_echo_b_value():
    """90, n gradient lobes, 180, n lobes, echo. A lobe is an interval of dt
    at the state's current order followed by the shift, so the dephasing half
    carries the constant-gradient (Stejskal-Tanner) b-value sum
    sum_{l<n} (l^2 + l + 1/3) = n^3 / 3; on the way back an F- state at order
    l is weighted l^2 + l + 1/3 as well (the reference applies one transverse
    weight to F+ and F-, src/torchsim/_epg_cpu.cpp:677-715)."""
    from torchsim_amd.sequence._events import Delay, Excitation, Readout, Refocusing, compose, pack

    n, dt, rate = 4, 0.005, 3.0
    lobe = [Delay(dt), Delay(0.0, action=16)]  # the interval, then SHIFT_AFTER
    parts = [Excitation(np.pi / 2, np.pi / 2)] + lobe * n + [Refocusing(np.pi, 0.0, crushed=False)]
    parts += lobe * n + [Readout(0.0)]
    events = oracle.events_of(pack(compose(*parts)[0]))
    damped = P.simulate(events, {"t1": 1e12, "t2": 1e12, "diffusion": rate}, states=16, outputs=1)
    clean = P.simulate(events, {"t1": 1e12, "t2": 1e12}, states=16, outputs=1)
    b = rate * float(events["duration"][2])
    back = sum(l * l + l + 1.0 / 3.0 for l in range(1, n + 1))
    ratio = abs(complex(damped[0, 0])) / abs(complex(clean[0, 0]))
    assert abs(abs(complex(clean[0, 0])) - 1.0) < 1e-9
    assert abs(ratio - np.exp(-b * (n**3 / 3 + back))) < 1e-12


def test_jacobian_matches_central_differences_with_exchange():
    events = _phased(_mrf_events(torch.tensor(_cases.benchmark_flips(40)), 10.0))
    tissue = {"t1": 900.0, "t2": 70.0, "fraction": 0.2, "exchange": 12.0, "t1b": 400.0, "t2b": 15.0,
              "shift": 20.0, "b1": 0.95}
    _, jac = P.jacobian(events, tissue, ("t1", "t2", "b1"), states=10, outputs=40)
    for k, (name, h) in enumerate((("t1", 0.5), ("t2", 0.05), ("b1", 1e-5))):
        up = dict(tissue, **{name: tissue[name] + h})
        dn = dict(tissue, **{name: tissue[name] - h})
        fd = (P.simulate(events, up, states=10, outputs=40) - P.simulate(events, dn, states=10, outputs=40)) / (2 * h)
        assert np.abs(jac[:, k].numpy() - fd.numpy()).max() / np.abs(fd.numpy()).max() < 1e-6
