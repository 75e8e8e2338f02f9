"""FSE (complex kernel + recovery post-processing) and balanced TrueFISP
(real subspace through half-turn signs; complex with an exchanging pool)
against the FP64 torch oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import packed as P
from tests import _cases

pytestmark = pytest.mark.gpu


def _fse_reference(seq_kwargs, t1, t2, m0, b1, names):
    from torchsim_amd.simulators import FSESimulator

    sim = FSESimulator()
    packed = sim.packed(**seq_kwargs)
    events = oracle.events_of(packed)
    echoes = len(np.atleast_1d(seq_kwargs["flip"]))
    left = seq_kwargs["TR"] - seq_kwargs["ESP"] * echoes
    base = {"t1": torch.tensor(t1, dtype=torch.float64), "t2": torch.tensor(t2, dtype=torch.float64),
            "m0": torch.tensor(m0, dtype=torch.float64), "b1": torch.tensor(b1, dtype=torch.float64)}

    def run(t1_, t2_, m0_, b1_):
        s = P.simulate(events, {"t1": t1_, "t2": t2_, "b1": b1_}, states=10, outputs=echoes)
        rec = torch.exp(-1e3 / t1_ * left * 1e-3)[:, None]
        return m0_[:, None] * s * (1 - rec) / (1 - rec * s)

    order = ("T1", "T2", "M0", "B1")
    primals = tuple(base[k] for k in ("t1", "t2", "m0", "b1"))
    out = run(*primals)
    columns = []
    for name in names:
        k = order.index(name)
        tangents = tuple(torch.ones_like(p) if j == k else torch.zeros_like(p) for j, p in enumerate(primals))
        columns.append(torch.func.jvp(run, primals, tangents)[1])
    return out.numpy(), torch.stack(columns, 1).numpy()


def _exact_half_turns(events):
    """The protocol's phases in FP64: a packed float32 phase within 1e-6 of a
    whole number of half turns is that multiple of pi exactly (the reference
    brings such pulses onto one axis exactly, sequence/_accelerators.py:500-571)."""
    phase = np.asarray(events["phase"], dtype=np.float64)
    turns = np.round(phase / np.pi)
    snapped = np.where(np.abs(phase - turns * np.pi) < 1e-6, turns * np.pi, phase)
    return dict(events, phase=snapped)


def test_fse_signal_and_jacobian():
    from torchsim_amd.simulators import FSESimulator

    flips = np.linspace(180, 120, 32)
    kw = dict(flip=torch.tensor(flips, dtype=torch.float32), ESP=5.0, TR=3000.0)
    t1, t2 = _cases.random_tissue(40, seed=31)
    m0 = np.linspace(0.5, 1.0, 40).astype(np.float32)
    b1 = np.linspace(0.8, 1.2, 40).astype(np.float32)
    names = ("T1", "T2", "M0", "B1")
    sig, jac = FSESimulator(**kw).jacobian(names, T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda(),
                                           M0=torch.tensor(m0).cuda(), B1=torch.tensor(b1).cuda())
    ref, rjac = _fse_reference(dict(flip=flips.astype(np.float32), ESP=5.0, TR=3000.0), t1, t2, m0, b1, names)
    assert _cases.rel_err(sig.cpu().numpy(), ref) < 1e-5
    for p in range(4):
        assert _cases.rel_err(jac.cpu().numpy()[:, p], rjac[:, p]) < 1e-4, names[p]
    only = FSESimulator(**kw).simulate(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda(),
                                       M0=torch.tensor(m0).cuda(), B1=torch.tensor(b1).cuda())
    assert torch.equal(only, sig)


def _reference_fp32_error(sim, t1, t2, ref):
    """The reference's own FP32 CPU kernel (built from source) on the same
    stream, brought onto one axis as the reference does; its error vs FP64."""
    from torchsim_amd.sequence import _engine

    if oracle.reference_module() is None:
        return None
    packed = sim.packed(**sim.bound)
    events = oracle.events_of(packed)
    _, flips, samples = _engine.half_turns(packed)
    events = dict(events, flip=(events["flip"] * flips.numpy()).astype(np.float32),
                  phase=np.zeros_like(events["phase"]))
    r32 = oracle.reference_real(events, {"t1": t1, "t2": t2}, states=1, outputs=ref.shape[-1], features=0)
    signs = np.ones(ref.shape[-1]) if samples is None else samples.numpy()
    return _cases.rel_err(1j * r32 * signs[None, :], ref)


@pytest.mark.parametrize("inversion", [None, "ideal"])
def test_truefisp_balanced_train(inversion):
    """A balanced train carries no shift, so a systematic error in the
    per-pulse coefficients is summed over the train's whole memory (T1/TR
    pulses). The kernels relax Z and apply the flip's versine as increments,
    which keeps the signal within the north-star 1e-5 of FP64; the
    reference's own FP32 kernel, where built, is reported beside it."""
    from torchsim_amd.simulators import TrueFISPSimulator

    t1, t2 = _cases.random_tissue(30, seed=32)
    sim = TrueFISPSimulator(flip=50.0, TR=4.5, nshots=300, inversion=inversion)
    sig, jac = sim.jacobian(("T1", "T2"), T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda())
    events = _exact_half_turns(oracle.events_of(sim.packed(**sim.bound)))
    ref, rjac = P.jacobian(events, {"t1": t1, "t2": t2}, ("t1", "t2"), states=1, outputs=300)
    ours = _cases.rel_err(sig.cpu().numpy(), ref.numpy())
    theirs = _reference_fp32_error(sim, t1, t2, ref.numpy())
    print(f"balanced train ({inversion}): ours {ours:.2e}, reference FP32 kernel {theirs}")
    assert ours < 1e-5
    for p in range(2):
        assert _cases.rel_err(jac.cpu().numpy()[:, p], rjac[:, p].numpy()) < 1e-4


def test_truefisp_with_exchanging_pool_and_off_resonance():
    from torchsim_amd.simulators import TrueFISPSimulator

    t1, t2 = _cases.random_tissue(20, seed=33)
    rng = np.random.default_rng(34)
    f = rng.uniform(0.05, 0.3, 20).astype(np.float32)
    k = rng.uniform(5, 20, 20).astype(np.float32)
    b0 = rng.uniform(-20, 20, 20).astype(np.float32)
    sim = TrueFISPSimulator(flip=40.0, TR=5.0, nshots=200, states=4)
    sig, jac = sim.jacobian(("T1", "T2"), T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda(),
                            B0=torch.tensor(b0).cuda(), pool_b_fraction=torch.tensor(f).cuda(),
                            pool_b_exchange=torch.tensor(k).cuda(), T1_pool_b=400.0, T2_pool_b=15.0)
    events = _exact_half_turns(oracle.events_of(sim.packed(**sim.bound)))
    tissue = {"t1": t1, "t2": t2, "b0": b0, "fraction": f, "exchange": k, "t1b": 400.0, "t2b": 15.0}
    ref, rjac = P.jacobian(events, tissue, ("t1", "t2"), states=4, outputs=200)
    assert _cases.rel_err(sig.cpu().numpy(), ref.numpy()) < 1e-5
    for p in range(2):
        assert _cases.rel_err(jac.cpu().numpy()[:, p], rjac[:, p].numpy()) < 1e-4


def test_reference_property_names_reach_the_complex_kernel():
    """The reference's public names (sequence/_parameters.py:234-252):
    ``B1phase`` turns every pulse of an MRF train by a per-voxel transmit
    phase, and the ``poolB_*`` names give the exchanging pool; both against
    the FP64 oracle, and ``poolB_*`` bitwise equal to the other spellings."""
    from torchsim_amd.simulators import MRFSimulator, TrueFISPSimulator

    t1, t2 = _cases.random_tissue(40, seed=35)
    rng = np.random.default_rng(36)
    phase = rng.uniform(-1.0, 1.0, 40).astype(np.float32)
    b0 = rng.uniform(-15, 15, 40).astype(np.float32)
    flip = _cases.mrf_flips()[:150]
    sim = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0, states=12)
    sig, jac = sim.jacobian(("T1", "T2"), T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda(),
                            B1phase=torch.tensor(phase).cuda(), B0=torch.tensor(b0).cuda())
    events = oracle.events_of(sim.packed(**sim.bound))
    ref, rjac = P.jacobian(events, {"t1": t1, "t2": t2, "b1_phase": phase, "b0": b0}, ("t1", "t2"), states=12,
                           outputs=150)
    ref, rjac = 1j * ref, 1j * rjac  # MRFSimulator's factor
    assert _cases.rel_err(sig.cpu().numpy(), ref.numpy()) < 1e-5
    for p in range(2):
        assert _cases.rel_err(jac.cpu().numpy()[:, p], rjac[:, p].numpy()) < 1e-4
    # one transmit phase on every pulse turns the transverse frame, so the
    # samples by e^{i phase} (F+ = -i e^{i phi} sin(alpha) Z after a pulse)
    plain = sim.simulate(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda(), B0=torch.tensor(b0).cuda())
    turned = plain.cpu().numpy() * np.exp(1j * phase)[:, None]
    assert _cases.rel_err(sig.cpu().numpy(), turned) < 1e-5

    f = rng.uniform(0.05, 0.3, 40).astype(np.float32)
    k = rng.uniform(5, 20, 40).astype(np.float32)
    bssfp = TrueFISPSimulator(flip=40.0, TR=5.0, nshots=100, states=4)
    common = {"T1": torch.tensor(t1).cuda(), "T2": torch.tensor(t2).cuda()}
    theirs = bssfp.simulate(**common, poolB_fraction=torch.tensor(f).cuda(), poolB_exchange=torch.tensor(k).cuda(),
                            poolB_T1=400.0, poolB_T2=15.0, poolB_shift=30.0)
    ours = bssfp.simulate(**common, pool_b_fraction=torch.tensor(f).cuda(), pool_b_exchange=torch.tensor(k).cuda(),
                          T1_pool_b=400.0, T2_pool_b=15.0, pool_b_shift=30.0)
    torch.testing.assert_close(theirs, ours, rtol=0, atol=0)
    ev = _exact_half_turns(oracle.events_of(bssfp.packed(**bssfp.bound)))
    want = P.simulate(ev, {"t1": t1, "t2": t2, "fraction": f, "exchange": k, "t1b": 400.0, "t2b": 15.0,
                           "shift": 30.0}, states=4, outputs=100)
    assert _cases.rel_err(theirs.cpu().numpy(), want.numpy()) < 1e-5


def test_bssfp_sim_shapes_values_and_jacobian():
    """bssfp_sim is the reference's closed form (reference
    tests/functionals/test_bssfp_func.py shapes): FP32 on the GPU against the
    FP64 fixed-point solve at TE = TR (oracle/bssfp.py), and its forward-mode
    columns against FP64 central differences of the same closed form."""
    from oracle import bssfp as oracle_bssfp
    from torchsim_amd import bssfp_sim
    from torchsim_amd.simulators.bssfp import steady_state

    flip = np.linspace(5.0, 60.0, 100, dtype=np.float32)
    dev = "cuda"
    assert bssfp_sim(5.0, TE=2.0, TR=10.0, T1=1000.0, T2=100.0, device=dev).shape == ()
    assert bssfp_sim(flip, TE=2.0, TR=10.0, T1=1000.0, T2=100.0, device=dev).shape == (100,)
    assert bssfp_sim(flip, TE=2.0, TR=10.0, T1=(200, 500, 1000.0), T2=100.0, device=dev).shape == (3, 100)
    assert bssfp_sim(flip, TE=2.0, TR=10.0, T1=1000.0, T2=100.0, diff="T1", device=dev)[1].shape == (100,)
    assert bssfp_sim(flip, TE=2.0, TR=10.0, T1=(200, 500, 1000.0), T2=100.0, diff="T1",
                     device=dev)[1].shape == (3, 100)
    assert bssfp_sim(flip, TE=2.0, TR=10.0, T1=1000.0, T2=100.0, diff=("T1", "T2"), device=dev)[1].shape == (2, 100)
    assert bssfp_sim(flip, TE=2.0, TR=10.0, T1=(200, 500, 1000.0), T2=100.0, diff=("T1", "T2"),
                     device=dev)[1].shape == (3, 2, 100)

    rng = np.random.default_rng(8)
    t1 = rng.uniform(300, 2500, 16).astype(np.float32)
    t2 = rng.uniform(20, 200, 16).astype(np.float32)
    b0 = rng.uniform(-60, 60, 16).astype(np.float32)
    tr = 6.0
    sig = bssfp_sim(flip, TE=tr, TR=tr, T1=torch.tensor(t1, device=dev), T2=torch.tensor(t2, device=dev),
                    B0=torch.tensor(b0, device=dev), phase_inc=180.0).cpu().numpy()
    ref = np.array([[oracle_bssfp.steady_state(float(t1[v]), float(t2[v]), 1.0, float(b0[v]), 0.0, float(f), tr, tr,
                                               180.0) for f in flip] for v in range(16)])
    err = _cases.rel_err(sig, ref)
    assert err < 1e-5, err

    names = ("T1", "T2", "B0")
    _, jac = bssfp_sim(flip, TE=2.5, TR=tr, T1=torch.tensor(t1, device=dev), T2=torch.tensor(t2, device=dev),
                       B0=torch.tensor(b0, device=dev), phase_inc=90.0, diff=names)
    jac = jac.cpu().numpy()
    base = {"T1": t1.astype(np.float64), "T2": t2.astype(np.float64), "B0": b0.astype(np.float64)}

    def closed(values):
        t = lambda a: torch.tensor(a, dtype=torch.float64)[:, None]  # noqa: E731
        return steady_state(t(values["T1"]), t(values["T2"]), torch.tensor(1.0, dtype=torch.float64), t(values["B0"]),
                            torch.tensor(0.0, dtype=torch.float64), torch.tensor(flip, dtype=torch.float64),
                            torch.tensor(tr, dtype=torch.float64), torch.tensor(2.5, dtype=torch.float64),
                            torch.tensor(90.0, dtype=torch.float64)).numpy()

    for p, name in enumerate(names):
        step = 1e-4 * np.maximum(np.abs(base[name]), 1.0)
        up = {**base, name: base[name] + step}
        down = {**base, name: base[name] - step}
        fd = (closed(up) - closed(down)) / (2 * step)[:, None]
        err = _cases.rel_err(jac[:, p], fd)
        assert err < 1e-4, (name, err)
