"""Pins the FP64 oracle before anything is checked against it.

Each check is chosen so a plausible mistake fails one of them:
- the reference's published benchmark checksums (benchmarks/results/
  torchsim-{forward,jacobian,jacobian1}-*.json), which also pin the packer;
- the reference's own FP32 CPU kernel built from source (oracle/_ref);
- an isochromat (Bloch) brute force at a state count that truncates nothing:
  pins the shift direction, the T-matrix and relaxation/recovery;
- central finite differences for every tangent direction;
- closed forms: first sample after an inversion, the Ernst steady state
  under ideal spoiling.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import isochromat
from tests import _cases


def _packed(flip, tr, **kw):
    from torchsim_amd.simulators import MRFSimulator

    return MRFSimulator().packed(flip=flip, TR=tr, **kw)


# (atoms, checksum) from benchmarks/results/torchsim-*-cpu-t4.json (CPU rows;
# forward-1000 from the t1 row). Task: benchmarks/_common.py:25-45, states 32.
FORWARD = {1: 47.452178955078125, 10: 219.9136505126953, 100: 2212.482177734375, 1000: 22137.09765625,
           10000: 221383.171875}
JACOBIAN = {1: 0.2923477292060852, 10: 1.5227932929992676, 100: 15.510462760925293, 1000: 155.29290771484375}
JACOBIAN_T1 = {1: -0.16639403998851776, 10: -0.5373333096504211, 100: -4.9750776290893555,
               1000: -49.40011978149414}


def _benchmark_events():
    flip = torch.tensor(_cases.benchmark_flips(500), dtype=torch.float32)
    return oracle.events_of(_packed(flip, 10.0))


def _grid(atoms):
    t1, t2 = _cases.benchmark_grid(atoms)
    # the benchmark hands torch float32 tissue to the simulator
    return t1.astype(np.float32).astype(np.float64), t2.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("atoms", sorted(FORWARD))
def test_forward_checksums(atoms):
    t1, t2 = _grid(atoms)
    signal, _ = oracle.real(_benchmark_events(), t1=t1, t2=t2, states=32, outputs=500)
    checksum = (-signal).sum()  # MRFSimulator returns 1j * (i m0 a0)
    assert abs(checksum - FORWARD[atoms]) / abs(FORWARD[atoms]) < 2e-6


@pytest.mark.parametrize("atoms", sorted(JACOBIAN))
def test_jacobian_checksums(atoms):
    t1, t2 = _grid(atoms)
    _, tangents = oracle.real(
        _benchmark_events(), t1=t1, t2=t2, states=32, outputs=500, tangents=[{"t1": 1.0}, {"t2": 1.0}]
    )
    checksum = (-tangents).sum()
    assert abs(checksum - JACOBIAN[atoms]) / abs(JACOBIAN[atoms]) < 2e-6
    _, tangents = oracle.real(_benchmark_events(), t1=t1, t2=t2, states=32, outputs=500, tangents=[{"t1": 1.0}])
    checksum = (-tangents).sum()
    assert abs(checksum - JACOBIAN_T1[atoms]) / abs(JACOBIAN_T1[atoms]) < 2e-6


def test_checksums_fail_with_a_planted_mistake():
    """A dropped recovery term in the oracle is caught by the checksums."""
    events = _benchmark_events()
    t1, t2 = _grid(10)
    good, _ = oracle.real(events, t1=t1, t2=t2, states=32, outputs=500)
    shifted = dict(events)
    shifted["duration"] = events["duration"] * (1 + 1e-3)  # a 0.1% clock error
    bad, _ = oracle.real(shifted, t1=t1, t2=t2, states=32, outputs=500)
    assert abs((-good).sum() - FORWARD[10]) / FORWARD[10] < 2e-6
    assert abs((-bad).sum() - FORWARD[10]) / FORWARD[10] > 1e-5


@pytest.mark.parametrize("b1,ie,m0", [(1.0, 1.0, 1.0), (0.8, 0.9, 0.6), (1.3, 0.95, 1.0)])
def test_isochromat_brute_force(b1, ie, m0):
    flip = torch.tensor(np.linspace(10, 70, 25))
    tr = torch.tensor(np.linspace(8, 14, 25))
    events = oracle.events_of(_packed(flip, tr, TI=5.0))
    signal, _ = oracle.real(events, t1=800.0, t2=60.0, b1=b1, ie=ie, m0=m0, states=64, outputs=25)
    samples = isochromat.simulate(events, t1=800.0, t2=60.0, b1=b1, ie=ie, m0=m0, isochromats=256)
    assert np.abs(samples.real).max() < 1e-12
    assert np.abs(samples.imag - signal[0]).max() < 1e-12


def test_truncation_differs_from_brute_force_when_orders_run_out():
    """The oracle truncates like the reference: with 4 orders it must NOT match."""
    flip = torch.tensor(np.linspace(10, 70, 25))
    events = oracle.events_of(_packed(flip, 10.0))
    signal, _ = oracle.real(events, t1=800.0, t2=200.0, states=4, outputs=25)
    samples = isochromat.simulate(events, t1=800.0, t2=200.0, isochromats=256)
    assert np.abs(samples.imag - signal[0]).max() > 1e-4


def test_first_sample_closed_form():
    flip = torch.tensor(np.array([37.0, 20.0, 20.0]))
    events = oracle.events_of(_packed(flip, 10.0))
    signal, _ = oracle.real(events, t1=900.0, t2=70.0, m0=0.7, ie=0.85, b1=1.1, states=10, outputs=3)
    alpha = float(events["flip"][1])  # the packed (float32) first flip
    assert abs(alpha - np.deg2rad(37.0)) < 1e-7
    assert abs(signal[0, 0] - 0.7 * 0.85 * np.sin(alpha * 1.1)) < 1e-15


def test_spoiled_steady_state_is_ernst():
    """A spoiled train at constant flip and TR settles at the Ernst signal."""
    from torchsim_amd.sequence._events import SPOILED, compose, pack
    from torchsim_amd.sequence._events import Excitation

    alpha, tr_s, t1 = np.deg2rad(15.0), 0.01, 700.0
    parts = []
    for _ in range(600):
        parts += [Excitation(alpha), SPOILED.readout(0.0, duration_s=tr_s)]
    events_raw, _ = compose(*parts)
    events = oracle.events_of(pack(events_raw))
    signal, _ = oracle.real(events, t1=t1, t2=50.0, states=4, outputs=600)
    # the packed (float32) flip and interval
    alpha, tr_s = float(events["flip"][0]), float(events["duration"][4])
    e1 = np.exp(-tr_s * 1000.0 / t1)
    ernst = np.sin(alpha) * (1 - e1) / (1 - np.cos(alpha) * e1)
    # From +z an excitation leaves a0 = -sin(alpha) Mz (recorded i m0 a0).
    assert abs(-signal[0, -1] - ernst) < 1e-12


@pytest.mark.parametrize("name,step", [("t1", 0.5), ("t2", 0.05), ("b1", 1e-5), ("ie", 1e-5), ("m0", 1e-5)])
def test_tangents_match_central_differences(name, step):
    flip = torch.tensor(_cases.benchmark_flips(120))
    events = oracle.events_of(_packed(flip, 9.0))
    base = {"t1": 1100.0, "t2": 90.0, "b1": 0.93, "ie": 0.9, "m0": 0.8}
    _, tangent = oracle.real(events, **base, states=12, outputs=120, tangents=[{name: 1.0}])
    plus = dict(base, **{name: base[name] + step})
    minus = dict(base, **{name: base[name] - step})
    fd = (oracle.real(events, **plus, states=12, outputs=120)[0] - oracle.real(events, **minus, states=12, outputs=120)[0]) / (2 * step)
    assert np.abs(tangent[:, 0] - fd).max() / np.abs(fd).max() < 1e-6


def test_event_tangents_match_central_differences():
    flip = torch.tensor(_cases.benchmark_flips(60))
    events = oracle.events_of(_packed(flip, 9.0))
    count = len(events["kind"])
    rng = np.random.default_rng(4)
    vflip = rng.normal(size=count)
    vdur = rng.normal(size=count) * 1e-3
    _, tangent = oracle.real(events, t1=900.0, t2=80.0, states=12, outputs=60,
                             tangents=[{"flip": vflip, "duration": vdur}])
    h = 1e-5
    up = dict(events, flip=events["flip"] + h * vflip, duration=events["duration"] + h * vdur)
    dn = dict(events, flip=events["flip"] - h * vflip, duration=events["duration"] - h * vdur)
    fd = (oracle.real(up, t1=900.0, t2=80.0, states=12, outputs=60)[0]
          - oracle.real(dn, t1=900.0, t2=80.0, states=12, outputs=60)[0]) / (2 * h)
    assert np.abs(tangent[:, 0] - fd).max() / np.abs(fd).max() < 1e-6


def test_reference_cpu_kernel_agrees():
    """The reference's FP32 CPU kernel, built from source, sits within FP32
    rounding of the FP64 oracle at config-0 shape (880 pulses, S 10)."""
    if oracle.reference_module() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    flip = torch.tensor(_cases.mrf_flips(), dtype=torch.float32)
    events = oracle.events_of(_packed(flip, 10.0))
    t1, t2 = _cases.random_tissue(200, seed=0)
    signal, tangents = oracle.real(events, t1=t1, t2=t2, states=10, outputs=880,
                                   tangents=[{"t1": 1.0}, {"t2": 1.0}])
    ref = oracle.reference_real(events, {"t1": t1, "t2": t2}, states=10, outputs=880, features=0)
    assert _cases.rel_err(ref, signal) < 1e-5
    for k, name in enumerate(("t1", "t2")):
        ref_t = oracle.reference_real(events, {"t1": t1, "t2": t2}, states=10, outputs=880, features=0,
                                      tangent={name: np.ones(200, dtype=np.float32)})
        assert _cases.rel_err(ref_t, tangents[:, k]) < 1e-4
