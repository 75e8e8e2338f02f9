"""Pins the FP64 complex oracle (oracle/packed.py)."""

from __future__ import annotations

import numpy as np
import pytest
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


def _fid(delays_s, tissue, *, plant=None):
    """A 90 deg pulse from equilibrium, then one sample after each delay."""
    from torchsim_amd.sequence._events import Delay, Excitation, Readout, compose, pack

    parts = [Excitation(np.pi / 2, 0.0)]
    last = 0.0
    for d in delays_s:
        parts += [Delay(d - last), Readout(0.0)]
        last = d
    events = oracle.events_of(pack(compose(*parts)[0]))
    return P.simulate(events, tissue, states=4, outputs=len(delays_s), _plant=plant).numpy()[0]


_POOLED = {"t1": 900.0, "t2": 70.0, "fraction": 0.2, "exchange": 15.0, "t1b": 400.0, "t2b": 15.0, "shift": 0.0}


@pytest.mark.parametrize("plant", [None, "swap_exchange"])
def test_exchange_equilibrium_is_stationary(plant):
    """With no pulse, populations (1 - f, f) are the fixed point of relaxation
    plus exchange (detailed balance k_ab (1 - f) = k_ba f with k_ab = k f,
    k_ba = k (1 - f); reference src/torchsim/_epg_cpu.cpp:895-896). A
    90-degree pulse read at once returns the total, 1, whatever waited
    before it. Swapping the two rates moves the fixed point, which this
    catches (the planted case must fail)."""
    from torchsim_amd.sequence._events import Delay, Excitation, Readout, compose, pack

    events = oracle.events_of(pack(compose(Delay(0.3), Delay(0.7), Excitation(np.pi / 2, 0.0), Readout(0.0))[0]))
    sample = P.simulate(events, _POOLED, states=4, outputs=1, _plant=plant).numpy()[0, 0]
    error = abs(sample - (-1j))
    if plant is None:
        assert error < 1e-12
    else:
        assert error > 1e-2


@pytest.mark.parametrize("plant", [None, "swap_exchange"])
def test_fast_exchange_coalesces_at_the_population_weighted_frequency(plant):
    """Two lines Δf apart (reference tests/sequence/test_exchange.py:173-194):
    with no exchange half a period of the shift leaves |(1 - f) - f| of the
    signal; past coalescence the lines merge into one at the population
    average f·Δf, so the magnitude survives and the phase turns by
    -2π f Δf t (the offset's sign convention, e^{-2πi Δf t}). A swapped
    exchange weights the average by 1 - f instead."""
    shift, f = -420.0, 0.2
    half = 0.5 / abs(shift)
    tissue = dict(_POOLED, t1=1e9, t2=1e9, t1b=1e9, t2b=1e9, shift=shift, fraction=f)
    slow = _fid((0.0, half), dict(tissue, exchange=0.0), plant=plant)
    fast = _fid((0.0, half), dict(tissue, exchange=2.0e5), plant=plant)
    assert abs(abs(slow[1] / slow[0]) - (1 - 2 * f)) < 1e-9
    assert abs(fast[1] / fast[0]) > 0.99
    turn = np.angle(fast[1] / fast[0])
    expected = np.angle(np.exp(-2j * np.pi * f * shift * half))
    if plant is None:
        assert abs(turn - expected) < 2e-2
    else:
        assert abs(turn - expected) > 0.5


def _reference_or_skip():
    if oracle.reference_module() is None:
        pytest.skip("oracle/_ref (the reference CPU kernel built from source) is not built")


def _pooled_tissue():
    base = {"t1": np.array([900.0, 1200.0, 600.0]), "t2": np.array([70.0, 90.0, 40.0]),
            "b1": np.array([0.95, 1.05, 0.8]), "b0": np.array([5.0, -3.0, 0.0])}
    pool = {"fraction": np.array([0.2, 0.1, 0.3]), "exchange": np.array([12.0, 5.0, 20.0]),
            "t1b": np.array([400.0, 350.0, 500.0]), "t2b": np.array([15.0, 12.0, 20.0]),
            "shift": np.array([30.0, -40.0, 10.0])}
    reference = dict(base, pool_b_fraction=pool["fraction"], pool_b_exchange=pool["exchange"],
                     t1_pool_b=pool["t1b"], t2_pool_b=pool["t2b"], pool_b_shift=pool["shift"])
    return {**base, **pool}, reference


_REF_NAME = {"t1": "t1", "t2": "t2", "b1": "b1", "b0": "b0", "fraction": "pool_b_fraction",
             "exchange": "pool_b_exchange", "t1b": "t1_pool_b", "t2b": "t2_pool_b", "shift": "pool_b_shift",
             "diffusion": "diffusion"}


@pytest.mark.parametrize("plant", [None, "swap_exchange"])
def test_exchanging_pool_matches_the_reference_kernel(plant):
    """The exchanging pool with relaxation on, f != 1/2, a chemical shift,
    off-resonance, B1 and phase-cycled pulses against the reference's own
    FP32 complex kernel (pool kind 2; src/torchsim/_epg_cpu.cpp:885-932,
    1714-1764, 4216-4566) built from source. FP32 vs FP64: 1e-5 of the
    largest sample. The planted rate swap is off by ~20%."""
    _reference_or_skip()
    events = _phased(_mrf_events(torch.tensor(_cases.benchmark_flips(80)), torch.tensor(np.linspace(8, 14, 80))))
    ours, reference = _pooled_tissue()
    ref = oracle.reference_complex(events, reference, states=10, outputs=80, exchanging=True)
    sig = P.simulate(events, ours, states=10, outputs=80, _plant=plant).numpy()
    error = _cases.rel_err(sig, ref)
    if plant is None:
        assert error < 1e-5
    else:
        assert error > 1e-2


@pytest.mark.parametrize("complex_path", [False, True])
@pytest.mark.parametrize("plant", [None, "z_transverse_weight"])
def test_diffusion_matches_the_reference_kernel(complex_path, plant):
    """Per-order diffusion damping (weight l^2 on Z, l^2 + l + 1/3 on F+ and
    F-; src/torchsim/_epg_cpu.cpp:466-474, 677-715) against the reference's
    FP32 kernel, on an MRF train whose crushers carry Z to orders above 0,
    in the real subspace and (phases, B0, B1) on the complex path. The
    planted transverse weight on Z is off by ~15%."""
    _reference_or_skip()
    flips = torch.tensor(_cases.benchmark_flips(80))
    events = _mrf_events(flips, 10.0)
    tissue = {"t1": np.array([900.0, 1200.0, 600.0]), "t2": np.array([70.0, 90.0, 40.0]),
              "diffusion": np.array([2.5, 1.0, 5.0])}
    if complex_path:
        events = _phased(events)
        tissue.update(b0=np.array([5.0, -3.0, 0.0]), b1=np.array([0.95, 1.05, 0.8]))
    ref = oracle.reference_complex(events, tissue, states=10, outputs=80)
    sig = P.simulate(events, tissue, states=10, outputs=80, _plant=plant).numpy()
    error = _cases.rel_err(sig, ref)
    if plant is None:
        assert error < 1e-5
    else:
        assert error > 1e-2


def test_c_oracle_diffusion_matches_the_reference_kernel():
    """The C oracle's diffusion row (real subspace) against the reference's
    real kernel and against the complex FP64 oracle."""
    _reference_or_skip()
    events = _mrf_events(torch.tensor(_cases.benchmark_flips(80)), 10.0)
    t1, t2 = np.array([900.0, 1200.0, 600.0]), np.array([70.0, 90.0, 40.0])
    rate = np.array([2.5, 1.0, 5.0])
    sig, tan = oracle.real(events, t1=t1, t2=t2, diffusion=rate, states=10, outputs=80,
                           tangents=[{"diffusion": 1.0}])
    ref = oracle.reference_real(events, {"t1": t1, "t2": t2, "diffusion": rate}, states=10, outputs=80,
                                features=oracle.FEATURE_ALL)
    assert _cases.rel_err(sig, ref) < 1e-5
    ref_t = oracle.reference_real(events, {"t1": t1, "t2": t2, "diffusion": rate}, states=10, outputs=80,
                                  features=oracle.FEATURE_ALL, tangent={"diffusion": 1.0})
    assert _cases.rel_err(tan[:, 0], ref_t) < 1e-4
    packed = P.simulate(events, {"t1": t1, "t2": t2, "diffusion": rate}, states=10, outputs=80).numpy()
    assert np.abs(packed.imag - sig).max() < 1e-13


def _balanced_events(pulses=60):
    from torchsim_amd.simulators import TrueFISPSimulator

    return oracle.events_of(TrueFISPSimulator().packed(flip=torch.full((pulses,), 40.0), TR=6.0, nshots=pulses))


@pytest.mark.parametrize("name", ["t1", "t2", "b1", "b0", "fraction", "exchange", "t1b", "t2b", "shift"])
@pytest.mark.parametrize("train", ["mrf", "balanced"])
def test_exchanging_jacobian_matches_the_reference_jvp(name, train):
    """Every tissue column of the exchanging pool (and B0) against the
    reference's forward-mode kernel (``simulate_jvp``, pool kind 2;
    src/torchsim/_epg_cpu.cpp:2469-3005, 11017-11148), 1e-4 of the column.
    Both trains: a spoiled MRF train's samples do not depend on B0 (every
    pathway that refocuses at order 0 spent as long dephased as rephased),
    so the B0 column is checked on the balanced train, where it does."""
    _reference_or_skip()
    if train == "mrf" and name == "b0":
        pytest.skip("B0 does not reach a spoiled train's samples")
    events = _phased(_mrf_events(torch.tensor(_cases.benchmark_flips(60)), 10.0)) if train == "mrf" else _balanced_events()
    ours, reference = _pooled_tissue()
    _, jac = P.jacobian(events, ours, (name,), states=10, outputs=60)
    ref = oracle.reference_complex(events, reference, states=10, outputs=60, exchanging=True,
                                   tangent={_REF_NAME[name]: 1.0})
    assert _cases.rel_err(jac[:, 0].numpy(), ref) < 1e-4


def test_diffusion_jacobian_matches_the_reference_jvp():
    """The diffusion column, with off-resonance and phases, against the
    reference's forward-mode kernel."""
    _reference_or_skip()
    events = _phased(_mrf_events(torch.tensor(_cases.benchmark_flips(60)), 10.0))
    tissue = {"t1": np.array([900.0, 1200.0]), "t2": np.array([70.0, 90.0]), "diffusion": np.array([2.5, 1.0]),
              "b0": np.array([4.0, -2.0])}
    _, jac = P.jacobian(events, tissue, ("diffusion",), states=10, outputs=60)
    ref = oracle.reference_complex(events, tissue, states=10, outputs=60, tangent={"diffusion": 1.0})
    assert _cases.rel_err(jac[:, 0].numpy(), ref) < 1e-4


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
