"""The allocation-free entry point (tsim_simulate_jacobian_noalloc) against
tsim_simulate_jacobian through the C ABI: bitwise equal on two calls in a
row, the second smaller (fewer atoms, events, states and columns), so the
second runs on what the first left in the shared device workspace."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from tests import _cases

pytestmark = pytest.mark.gpu


def _call(entry, packed, tissue, *, states, columns, signs=None):
    from torchsim_amd import _native

    device = torch.device("cuda", 0)
    events = [value.to(device).contiguous() for value in packed.buffers()]
    atoms, outputs = tissue.atoms, packed.output_count
    signal = torch.full((atoms, outputs), float("nan"), dtype=torch.complex64, device=device)
    jacobian = torch.full((atoms, len(columns), outputs), float("nan"), dtype=torch.complex64, device=device)
    cols = (ctypes.c_int32 * len(columns))(*columns)
    ptrs = _native.pointers([*tissue.buffers, *events])
    status = getattr(_native.load(), entry)(
        ptrs, atoms, packed.event_count, states, outputs, tissue.features, tissue.stride, cols, len(columns),
        0.0, 1.0, None if signs is None else signs.data_ptr(), signal.data_ptr(), jacobian.data_ptr(), -1,
        _native.stream_of(device),
    )
    _native.check(status)
    torch.cuda.synchronize()
    return signal, jacobian


def test_noalloc_equals_the_allocating_entry_point_twice():
    from torchsim_amd.sequence import _engine
    from torchsim_amd.simulators import MRFSimulator

    device = torch.device("cuda", 0)
    cases = [
        # 880-pulse MRF dictionary slice with a B1 map: S 20, columns T1 T2 B1
        dict(pulses=880, atoms=4099, states=20, columns=(0, 1, 3), b1=True, seed=61),
        # smaller: 300 pulses, fewer atoms, S 10, columns T1 T2
        dict(pulses=300, atoms=1001, states=10, columns=(0, 1), b1=False, seed=62),
    ]
    from torchsim_amd import _native

    assert packed_events_limit() == _native.load().tsim_noalloc_max_events()
    for case in cases:
        print(f"noalloc case: {case}", flush=True)
        flips = torch.tensor(_cases.mrf_flips()[: case["pulses"]], dtype=torch.float32)
        packed = MRFSimulator().packed(flip=flips, TR=10.0)
        assert packed.event_count <= packed_events_limit()
        t1, t2 = _cases.random_tissue(case["atoms"], seed=case["seed"])
        props = {"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2)}
        if case["b1"]:
            props["b1"] = torch.tensor(np.random.default_rng(case["seed"]).uniform(0.6, 1.4, case["atoms"]),
                                       dtype=torch.float32)
        tissue = _engine.prepare_tissue(props, device)
        want = _call("tsim_simulate_jacobian", packed, tissue, states=case["states"], columns=case["columns"])
        got = _call("tsim_simulate_jacobian_noalloc", packed, tissue, states=case["states"], columns=case["columns"])
        assert torch.equal(got[0], want[0])
        assert torch.equal(got[1], want[1])
        assert bool(torch.isfinite(torch.view_as_real(got[0])).all())


def packed_events_limit() -> int:
    return 16384
