"""GPU parity of the fused real-subspace signal + Jacobian kernel against the
FP64 oracle (signal within 1e-5, Jacobian within 1e-4, relative to the
tensor's largest magnitude), through the public API and the C ABI."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from tests import _cases

pytestmark = pytest.mark.gpu

SIGNAL_TOL = 1e-5
JACOBIAN_TOL = 1e-4


def _oracle_mrf(flip_deg, tr, t1, t2, *, states, b1=None, m0=None, ie=None, names=()):
    from torchsim_amd.simulators import MRFSimulator

    packed = MRFSimulator().packed(flip=torch.tensor(flip_deg, dtype=torch.float32), TR=tr)
    rows = {"T1": "t1", "T2": "t2", "B1": "b1", "M0": "m0", "inv_efficiency": "ie"}
    signal, tangents = oracle.real(
        oracle.events_of(packed),
        t1=t1,
        t2=t2,
        b1=b1,
        m0=m0,
        ie=ie,
        states=states,
        outputs=packed.output_count,
        tangents=[{rows[name]: 1.0} for name in names],
    )
    # MRFSimulator returns 1j * (i m0 a0) = -m0 a0.
    return -signal, None if tangents is None else -tangents


def test_config0_mrf_sim_jacobian_t1_t2():
    from torchsim_amd import mrf_sim

    flip = _cases.mrf_flips()
    t1, t2 = _cases.random_tissue(1000, seed=0)
    sig, jac = mrf_sim(
        flip=torch.tensor(flip, dtype=torch.float32),
        TR=10.0,
        T1=torch.tensor(t1),
        T2=torch.tensor(t2),
        diff=("T1", "T2"),
        device="cuda",
    )
    assert sig.shape == (1000, 880) and jac.shape == (1000, 2, 880)
    assert sig.dtype == torch.complex64 and jac.dtype == torch.complex64
    ref_sig, ref_jac = _oracle_mrf(flip, 10.0, t1, t2, states=10, names=("T1", "T2"))
    sig = sig.cpu().numpy()
    jac = jac.cpu().numpy()
    assert np.abs(sig.imag).max() == 0.0
    assert _cases.rel_err(sig.real, ref_sig) < SIGNAL_TOL
    for p in range(2):
        assert _cases.rel_err(jac[:, p].real, ref_jac[:, p]) < JACOBIAN_TOL


def test_config1_dictionary_sample_t1_t2_b1():
    from torchsim_amd.simulators import MRFSimulator

    flip = _cases.mrf_flips()
    T1, T2, B1 = _cases.dictionary_grid()
    pick = np.random.default_rng(1).choice(T1.size, 1500, replace=False)
    t1, t2, b1 = T1[pick], T2[pick], B1[pick]
    seq = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0, states=20)
    sig, jac = seq.jacobian(
        ("T1", "T2", "B1"), T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda(), B1=torch.tensor(b1).cuda()
    )
    ref_sig, ref_jac = _oracle_mrf(flip, 10.0, t1, t2, b1=b1, states=20, names=("T1", "T2", "B1"))
    assert _cases.rel_err(sig.cpu().numpy().real, ref_sig) < SIGNAL_TOL
    jac = jac.cpu().numpy().real
    for p in range(3):
        assert _cases.rel_err(jac[:, p], ref_jac[:, p]) < JACOBIAN_TOL


@pytest.mark.parametrize("states", [1, 3, 7, 10, 13, 20, 32, 47, 64])
def test_state_counts_including_padded_tiles(states):
    from torchsim_amd.simulators import MRFSimulator

    flip = _cases.benchmark_flips(200)
    t1, t2 = _cases.random_tissue(96, seed=states)
    seq = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0, states=states)
    sig, jac = seq.jacobian(("T1", "T2"), T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda())
    ref_sig, ref_jac = _oracle_mrf(flip, 10.0, t1, t2, states=states, names=("T1", "T2"))
    assert _cases.rel_err(sig.cpu().numpy().real, ref_sig) < SIGNAL_TOL
    for p in range(2):
        assert _cases.rel_err(jac.cpu().numpy()[:, p].real, ref_jac[:, p]) < JACOBIAN_TOL


def test_every_tile_agrees_with_oracle():
    from torchsim_amd import _native
    from torchsim_amd.simulators import MRFSimulator

    flip = _cases.benchmark_flips(150)
    t1, t2 = _cases.random_tissue(70, seed=5)
    b1 = np.linspace(0.7, 1.3, 70).astype(np.float32)
    for tile, (sl, lanes) in enumerate(_native.tiles()):
        states = min(sl * lanes, 64)
        seq = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=9.0, states=states)
        sig, jac = seq.jacobian(
            ("T1", "T2", "B1"),
            T1=torch.tensor(t1).cuda(),
            T2=torch.tensor(t2).cuda(),
            B1=torch.tensor(b1).cuda(),
            tile=tile,
        )
        ref_sig, ref_jac = _oracle_mrf(flip, 9.0, t1, t2, b1=b1, states=states, names=("T1", "T2", "B1"))
        assert _cases.rel_err(sig.cpu().numpy().real, ref_sig) < SIGNAL_TOL, (tile, sl, lanes)
        for p in range(3):
            assert _cases.rel_err(jac.cpu().numpy()[:, p].real, ref_jac[:, p]) < JACOBIAN_TOL, (tile, p)


def test_m0_and_inversion_efficiency_columns():
    from torchsim_amd.simulators import MRFSimulator

    flip = _cases.benchmark_flips(120)
    t1, t2 = _cases.random_tissue(40, seed=3)
    m0 = np.linspace(0.5, 1.0, 40).astype(np.float32)
    ie = np.linspace(0.8, 1.0, 40).astype(np.float32)
    seq = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=12.0, states=10)
    names = ("T1", "M0", "inv_efficiency", "T2")
    sig, jac = seq.jacobian(
        names,
        T1=torch.tensor(t1).cuda(),
        T2=torch.tensor(t2).cuda(),
        M0=torch.tensor(m0).cuda(),
        inv_efficiency=torch.tensor(ie).cuda(),
    )
    ref_sig, ref_jac = _oracle_mrf(flip, 12.0, t1, t2, m0=m0, ie=ie, states=10, names=names)
    assert _cases.rel_err(sig.cpu().numpy().real, ref_sig) < SIGNAL_TOL
    for p in range(len(names)):
        assert _cases.rel_err(jac.cpu().numpy()[:, p].real, ref_jac[:, p]) < JACOBIAN_TOL, names[p]


def test_scalar_call_and_single_name_shapes():
    from torchsim_amd import mrf_sim

    flip = _cases.benchmark_flips(50)
    sig = mrf_sim(flip=flip, TR=10.0, T1=1000.0, T2=80.0, device="cuda")
    assert isinstance(sig, np.ndarray) and sig.shape == (50,)
    sig, jac = mrf_sim(flip=flip, TR=10.0, T1=np.array([1000.0, 800.0]), T2=80.0, diff="T1", device="cuda")
    assert sig.shape == (2, 50) and jac.shape == (2, 50)
    # First sample after an ideal inversion with TI = 0: -M0 * eta * sin(alpha_0).
    sig = mrf_sim(flip=flip, TR=10.0, T1=1000.0, T2=80.0, inv_efficiency=0.9, M0=0.7, device="cuda")
    expected = -0.7 * 0.9 * np.sin(np.deg2rad(np.float32(flip[0])))
    assert abs(sig[0].real - expected) < 1e-6


def test_phase_half_turns_flip_sample_signs():
    """Pulses a half turn apart stay in the real subspace through signs."""
    from torchsim_amd.simulators import MRFSimulator

    flip = _cases.benchmark_flips(100)
    phases = np.where(np.arange(100) % 2 == 0, 0.0, 180.0)
    t1, t2 = _cases.random_tissue(30, seed=9)
    seq = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0, phases=phases, states=10)
    sig = seq.simulate(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda()).cpu().numpy()
    # The same train at phase 0 with the odd flips negated and the odd samples negated.
    from torchsim_amd.simulators import MRFSimulator as M

    signed = np.where(np.arange(100) % 2 == 0, 1.0, -1.0)
    packed = M().packed(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0)
    events = oracle.events_of(packed)
    turning = (events["kind"] == 1) & ((events["action"] & 4) == 0)
    events["flip"] = events["flip"].copy()
    events["flip"][turning] *= signed
    ref, _ = oracle.real(events, t1=t1, t2=t2, states=10, outputs=100)
    ref = -ref * signed[None, :]
    assert _cases.rel_err(sig.real, ref) < SIGNAL_TOL


def test_same_alone_and_in_the_mix_bitwise():
    from torchsim_amd.simulators import MRFSimulator

    flip = _cases.benchmark_flips(200)
    t1, t2 = _cases.random_tissue(500, seed=11)
    seq = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0, states=20)
    sig, jac = seq.jacobian(("T1", "T2"), T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda())
    part = slice(123, 161)
    sig_a, jac_a = seq.jacobian(("T1", "T2"), T1=torch.tensor(t1[part]).cuda(), T2=torch.tensor(t2[part]).cuda())
    assert torch.equal(sig[part], sig_a)
    assert torch.equal(jac[part], jac_a)


def test_empty_atoms():
    from torchsim_amd.simulators import MRFSimulator

    seq = MRFSimulator(flip=torch.tensor(_cases.benchmark_flips(20), dtype=torch.float32), TR=10.0)
    sig, jac = seq.jacobian(("T1", "T2"), T1=torch.empty(0).cuda(), T2=torch.empty(0).cuda())
    assert sig.shape == (0, 20) and jac.shape == (0, 2, 20)


def test_reference_plane_entry_point_matches_reference_layout():
    """tsim_simulate writes the reference's (real, imag) planes: (0, m0 a0)."""
    from torchsim_amd import _native
    from torchsim_amd.sequence import _engine
    from torchsim_amd.simulators import MRFSimulator

    flip = _cases.benchmark_flips(60)
    t1, t2 = _cases.random_tissue(33, seed=2)
    packed = MRFSimulator().packed(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0)
    tissue = _engine.prepare_tissue({"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2)}, torch.device("cuda", 0))
    real = torch.empty((33, 60), device="cuda")
    imag = torch.empty((33, 60), device="cuda")
    events = [value.cuda().contiguous() for value in packed.buffers()]
    ptrs = _native.pointers([*tissue.buffers, *events, real, imag])
    lib = _native.load()
    status = lib.tsim_simulate(ptrs, 28, 33, 1, packed.event_count, 10, 60, 1, 1, 0, tissue.features, tissue.stride,
                               _native.stream_of(torch.device("cuda", 0)))
    _native.check(status)
    ref, _ = oracle.real(oracle.events_of(packed), t1=t1, t2=t2, states=10, outputs=60)
    assert float(real.abs().max()) == 0.0
    assert _cases.rel_err(imag.cpu().numpy(), ref) < SIGNAL_TOL


def test_staged_and_direct_output_writes_agree_bitwise():
    """Records that fill the outputs in order 0, 1, 2, ... are written through
    the kernel's LDS stage (K samples of a row at a time); a stream whose
    records fill them in reverse takes the direct per-sample stores. The
    arithmetic is the same, so after undoing the permutation the two agree
    bit for bit, and both match the oracle. 203 outputs: several full stages
    and a ragged last one."""
    import dataclasses

    from torchsim_amd.sequence import _engine
    from torchsim_amd.simulators import MRFSimulator

    pulses = 203
    flip = _cases.benchmark_flips(pulses)
    t1, t2 = _cases.random_tissue(77, seed=13)
    b1 = np.linspace(0.7, 1.3, 77).astype(np.float32)
    packed = MRFSimulator().packed(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0)
    recorded = packed.output_index >= 0
    reverse = packed.output_index.clone()
    reverse[recorded] = (pulses - 1) - packed.output_index[recorded]
    backwards = dataclasses.replace(packed, output_index=reverse)
    device = torch.device("cuda", 0)
    names = ("t1_ms", "t2_ms", "b1", "m0")
    tissue = _engine.prepare_tissue({"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2), "b1": torch.tensor(b1),
                                     "m0": torch.ones(77)}, device, differentiated=names)
    sig, jac = _engine.simulate_real(packed, tissue, states=20, columns=names, factor=1j)
    sig_r, jac_r = _engine.simulate_real(backwards, tissue, states=20, columns=names, factor=1j)
    assert torch.equal(sig, torch.flip(sig_r, dims=(-1,)))
    assert torch.equal(jac, torch.flip(jac_r, dims=(-1,)))
    ref, _ = _oracle_mrf(flip, 10.0, t1, t2, states=20, b1=b1)
    assert _cases.rel_err(sig.cpu().numpy(), ref) < SIGNAL_TOL


def test_every_covering_tile_gives_the_same_bits():
    """The tuner may pick any covering tile: each runs the same per-order
    arithmetic (padded tiles only keep their extra orders at zero), so the
    outputs agree bit for bit whatever the choice."""
    from torchsim_amd import _native
    from torchsim_amd.simulators import MRFSimulator

    flip = _cases.benchmark_flips(120)
    t1, t2 = _cases.random_tissue(300, seed=21)
    b1 = np.linspace(0.7, 1.3, 300).astype(np.float32)
    seq = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0, states=20)
    names = ("T1", "T2", "B1")
    args = dict(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda(), B1=torch.tensor(b1).cuda())
    shapes = _native.tiles()
    covering = [index for index, (sl, lanes) in enumerate(shapes) if sl * lanes >= 20]
    first = None
    for index in covering:
        sig, jac = seq.jacobian(names, tile=index, **args)
        if first is None:
            first = (sig, jac)
        else:
            assert torch.equal(sig, first[0]), shapes[index]
            assert torch.equal(jac, first[1]), shapes[index]


def test_tile_tuner_measures_once_per_configuration():
    from torchsim_amd.sequence import _tuning

    _tuning.forget()
    calls = []

    def probe(tile):
        calls.append(tile)
        torch.cuda._sleep(1000)

    device = torch.device("cuda", 0)
    candidates = _tuning.candidates(20)
    assert candidates and all(sl * lanes >= 20 for sl, lanes in (_native_tiles()[i] for i in candidates))
    chosen = _tuning.tile_for(device, 20, ("t1_ms",), 0, probe)
    assert chosen in candidates
    measured = len(calls)
    assert measured >= len(candidates)
    assert _tuning.tile_for(device, 20, ("t1_ms",), 0, probe) == chosen
    assert len(calls) == measured  # cached: no second measurement
    _tuning.forget()


def _native_tiles():
    from torchsim_amd import _native

    return _native.tiles()


@pytest.mark.parametrize("atoms", [16383, 1000])
def test_launch_shape_switch_leaves_every_atom_bitwise(atoms):
    """Launches of fewer than 65,536 lanes run one-wave blocks, larger ones
    four-wave blocks: the same atoms give the same bits either way (tile
    (5, 4): 16,384 atoms is the switch)."""
    from torchsim_amd import _native
    from torchsim_amd.simulators import MRFSimulator

    print(f"case atoms={atoms}", flush=True)
    flip = torch.tensor(_cases.mrf_flips()[:200], dtype=torch.float32)
    seq = MRFSimulator(flip=flip, TR=10.0, states=20)
    t1, t2 = _cases.random_tissue(16385, seed=77)
    T1, T2 = torch.tensor(t1).cuda(), torch.tensor(t2).cuda()
    tile = _native.tiles().index((5, 4))
    small, jsmall = seq.jacobian(("T1", "T2"), T1=T1[:atoms], T2=T2[:atoms], tile=tile)
    large, jlarge = seq.jacobian(("T1", "T2"), T1=T1, T2=T2, tile=tile)
    torch.testing.assert_close(small, large[:atoms], rtol=0, atol=0)
    torch.testing.assert_close(jsmall, jlarge[:atoms], rtol=0, atol=0)
