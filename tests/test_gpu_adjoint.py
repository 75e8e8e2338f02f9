"""Sequence-design gradients through the checkpointed adjoint kernel.

Loss = sum over atoms of tr(F^-1) of (T1, T2, M0) at noise sigma = 1; the
gradient w.r.t. every flip (deg) and TR (ms) comes from
``crlb(J).sum().backward()`` through the fused Jacobian's autograd Function.
Checked against (a) the FP64 torch oracle's own autograd gradient (full
vector, reduced size) and (b) FP64 central differences of the C oracle's loss
along random directions (full config: 1024 atoms x 880 pulses), plus exact
invariants (the last TR is never played, checkpoint length changes nothing).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import packed as P
from tests import _cases

pytestmark = pytest.mark.gpu
# max|g - g_ref| / max|g_ref|. The FP32 adjoint meets the FP64 gradient and
# the reference's FP32 gradient to ~1e-6 (flips 2.7e-7, TRs 1.0e-6 at the
# full design shape); a dropped term or a wrong sign is off by far more.
GRAD_TOL = 1e-5
REF_GRAD_TOL = 1e-5


def _design_inputs(pulses, atoms, seed=1):
    rng = np.random.default_rng(seed)
    flips = rng.uniform(2.0, 70.0, pulses)
    trs = rng.uniform(8.0, 15.0, pulses)
    t1, t2 = _cases.random_tissue(atoms, seed=0)
    return flips, trs, t1, t2


def _gpu_loss_and_grad(flips, trs, t1, t2, states=10, checkpoint_pulses=64):
    from torchsim_amd.optim import crlb
    from torchsim_amd.simulators import MRFSimulator

    flip = torch.tensor(flips, dtype=torch.float32, requires_grad=True)
    tr = torch.tensor(trs, dtype=torch.float32, requires_grad=True)
    seq = MRFSimulator(states=states, checkpoint_pulses=checkpoint_pulses)
    _, jac = seq.jacobian(("T1", "T2", "M0"), flip=flip, TR=tr, T1=torch.tensor(t1).cuda(),
                          T2=torch.tensor(t2).cuda(), M0=torch.ones(len(t1)).cuda())
    loss = crlb(jac).sum()
    loss.backward()
    return float(loss), flip.grad.numpy().astype(np.float64), tr.grad.numpy().astype(np.float64)


def _oracle_events(flips_deg, trs_ms):
    """The packed MRF stream in FP64 built straight from the protocol."""
    from torchsim_amd.simulators import MRFSimulator

    packed = MRFSimulator().packed(flip=torch.tensor(flips_deg), TR=torch.tensor(trs_ms))
    events = oracle.events_of(packed)
    kind, action = events["kind"], events["action"]
    rf = np.flatnonzero((kind == 1) & ((action & 4) == 0))
    flip = np.zeros(len(kind))
    flip[rf] = np.deg2rad(flips_deg)
    duration = np.zeros(len(kind))
    duration[rf[1:]] = np.asarray(trs_ms[:-1]) * 1e-3
    events = dict(events, flip=flip, duration=duration)
    return events


def _oracle_loss(flips_deg, trs_ms, t1, t2, states=10):
    events = _oracle_events(flips_deg, trs_ms)
    n = len(flips_deg)
    _, tan = oracle.real(events, t1=t1, t2=t2, m0=np.ones(len(t1)), states=states, outputs=n,
                         tangents=[{"t1": 1.0}, {"t2": 1.0}, {"m0": 1.0}])
    jac = -tan  # MRFSimulator: 1j * (i m0 a0)
    fisher = jac @ np.transpose(jac, (0, 2, 1))
    return float(np.trace(np.linalg.inv(fisher), axis1=1, axis2=2).sum())


def test_gradient_matches_fp64_autograd_reduced():
    flips, trs, t1, t2 = _design_inputs(60, 3)
    loss, gf, gt = _gpu_loss_and_grad(flips, trs, t1, t2)
    flip = torch.tensor(flips, dtype=torch.float64, requires_grad=True)
    tr = torch.tensor(trs, dtype=torch.float64, requires_grad=True)
    from torchsim_amd.simulators import MRFSimulator

    packed = MRFSimulator().packed(flip=flip, TR=tr)
    events = {"duration": packed.duration.double(), "flip": packed.flip.double(), "phase": packed.phase.detach(),
              "kind": packed.kind, "action": packed.action, "output_index": packed.output_index}
    base = (torch.tensor(t1, dtype=torch.float64), torch.tensor(t2, dtype=torch.float64), torch.ones(3, dtype=torch.float64))

    def run(a, b, c):
        return P.simulate(events, {"t1": a, "t2": b, "m0": c}, states=10, outputs=60)

    cols = []
    for k in range(3):
        tangents = tuple(torch.ones(3, dtype=torch.float64) if j == k else torch.zeros(3, dtype=torch.float64) for j in range(3))
        cols.append(1j * torch.func.jvp(run, base, tangents)[1])
    jac = torch.stack(cols, 1)
    rows = torch.cat((jac.real, jac.imag), -1)
    ref_loss = torch.diagonal(torch.linalg.inv(rows @ rows.transpose(-1, -2)), dim1=-2, dim2=-1).sum()
    ref_loss.backward()
    e_flip, e_tr = _cases.rel_err(gf, flip.grad.numpy()), _cases.rel_err(gt, tr.grad.numpy())
    print(f"reduced design gradient vs FP64 autograd: flips {e_flip:.2e}, TRs {e_tr:.2e}")
    assert abs(loss - float(ref_loss.detach())) / abs(float(ref_loss.detach())) < 1e-5
    assert e_flip < GRAD_TOL
    assert e_tr < GRAD_TOL
    assert gt[-1] == 0.0  # the last TR is never played


def test_config2_gradient_vs_fp64_central_differences():
    flips, trs, t1, t2 = _design_inputs(880, 1024)
    loss, gf, gt = _gpu_loss_and_grad(flips, trs, t1, t2)
    ref = _oracle_loss(flips, trs, t1, t2)
    assert abs(loss - ref) / abs(ref) < 1e-5
    assert gt[-1] == 0.0
    rng = np.random.default_rng(3)
    h = np.rad2deg(1e-4)  # 1e-4 rad in the flip angle
    for trial in range(3):
        vf = rng.normal(size=880)
        vt = rng.normal(size=880)
        vf /= np.linalg.norm(vf)
        vt /= np.linalg.norm(vt)
        up = _oracle_loss(flips + h * vf, trs + h * vt, t1, t2)
        dn = _oracle_loss(flips - h * vf, trs - h * vt, t1, t2)
        fd = (up - dn) / (2 * h)
        ours = float(gf @ vf + gt @ vt)
        print(f"direction {trial}: adjoint {ours:.9e}, FP64 central difference {fd:.9e}")
        assert abs(ours - fd) / max(abs(fd), 1e-12) < 1e-4, (trial, ours, fd)


def test_config2_gradient_matches_the_reference_adjoint():
    """The full design shape (1,024 atoms x 880 pulses, S = 10): every flip
    and TR component against the reference's own gradient -- its FP32
    kernels driven as its autograd drives them (forward-over-reverse per
    tissue column, tests/golden/make_design_gradient.py), recomputed where
    oracle/_ref is built, else read from the committed fixture."""
    from pathlib import Path

    from tests.golden.make_design_gradient import reference_gradient

    flips, trs, t1, t2 = _design_inputs(880, 1024)
    loss, gf, gt = _gpu_loss_and_grad(flips, trs, t1, t2)
    if oracle.reference_module() is not None:
        ref_loss, ref_flip, ref_tr = reference_gradient(flips, trs, t1, t2, threads=16)
    else:
        golden = np.load(Path(__file__).resolve().parent / "golden" / "design_gradient_1024x880.npz")
        ref_loss, ref_flip, ref_tr = float(golden["loss"]), golden["grad_flip"], golden["grad_tr"]
    e_loss = abs(loss - ref_loss) / ref_loss
    e_flip, e_tr = _cases.rel_err(gf, ref_flip), _cases.rel_err(gt, ref_tr)
    print(f"config-2 gradient vs the reference's: loss {e_loss:.2e}, flips {e_flip:.2e}, TRs {e_tr:.2e}")
    assert e_loss < 1e-5
    assert e_flip < REF_GRAD_TOL and e_tr < REF_GRAD_TOL


def test_checkpoint_length_changes_nothing():
    flips, trs, t1, t2 = _design_inputs(150, 64)
    a = _gpu_loss_and_grad(flips, trs, t1, t2, checkpoint_pulses=64)
    b = _gpu_loss_and_grad(flips, trs, t1, t2, checkpoint_pulses=9)
    c = _gpu_loss_and_grad(flips, trs, t1, t2, checkpoint_pulses=100000)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[1], c[1])
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[2], c[2])


def test_signal_only_loss_first_order_adjoint():
    from torchsim_amd.simulators import MRFSimulator

    flips, trs, t1, t2 = _design_inputs(120, 32, seed=4)
    flip = torch.tensor(flips, dtype=torch.float32, requires_grad=True)
    tr = torch.tensor(trs, dtype=torch.float32, requires_grad=True)
    sig = MRFSimulator(states=12).simulate(flip=flip, TR=tr, T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda())
    loss = (sig.abs() ** 2).sum()
    loss.backward()

    def oracle_loss(fd, td):
        events = _oracle_events(fd, td)
        s, _ = oracle.real(events, t1=t1, t2=t2, states=12, outputs=120)
        return float((s**2).sum())

    rng = np.random.default_rng(5)
    h = 1e-3
    for _ in range(2):
        vf, vt = rng.normal(size=120), rng.normal(size=120)
        fd = (oracle_loss(flips + h * vf, trs + h * vt) - oracle_loss(flips - h * vf, trs - h * vt)) / (2 * h)
        ours = float(flip.grad.double() @ torch.tensor(vf) + tr.grad.double() @ torch.tensor(vt))
        assert abs(ours - fd) / abs(fd) < 1e-3


def test_sequence_design_reduces_the_bound():
    from torchsim_amd.optim import Bounded, SequenceDesign, crlb
    from torchsim_amd.simulators import MRFSimulator

    flips, trs, t1, t2 = _design_inputs(100, 16, seed=6)
    T1, T2 = torch.tensor(t1).cuda(), torch.tensor(t2).cuda()
    seq = MRFSimulator(states=10)

    def cost(flip, TR):
        _, jac = seq.jacobian(("T1", "T2", "M0"), flip=flip, TR=TR, T1=T1, T2=T2, M0=torch.ones(16).cuda())
        return (crlb(jac) / torch.stack((T1**2, T2**2, torch.ones_like(T1)), -1)).sum()

    design = SequenceDesign(cost, flip=Bounded(torch.tensor(flips, dtype=torch.float32), 1.0, 80.0),
                            TR=Bounded(torch.tensor(trs, dtype=torch.float32), 7.0, 16.0))
    result = design.minimize(iterations=8, learning_rate=0.05)
    assert float(result.loss[-1]) < float(result.loss[0])
    assert float(result.parameters["flip"].min()) >= 1.0 and float(result.parameters["flip"].max()) <= 80.0
