"""Pins the reference-side adjoint driver (oracle.reference_vjp,
oracle.reference_crlb_gradient) and the design-gradient fixture it wrote.

The driver runs the reference's own FP32 kernels (built from source into
oracle/_ref); these checks hold it to the FP64 torch oracle's autograd, so a
wrong cotangent plane, sign, column order or event mapping fails here."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from oracle import packed as P
from tests import _cases
from tests.golden.make_design_gradient import design_inputs, reference_gradient

GOLDEN = Path(__file__).resolve().parent / "golden" / "design_gradient_1024x880.npz"


def _reference_or_skip():
    if oracle.reference_module() is None:
        pytest.skip("oracle/_ref (the reference CPU kernel built from source) is not built")


def _fp64_design_gradient(flips, trs, t1, t2, states=10):
    from torchsim_amd.simulators import MRFSimulator

    atoms, pulses = len(t1), len(flips)
    flip = torch.tensor(flips, dtype=torch.float64, requires_grad=True)
    tr = torch.tensor(trs, dtype=torch.float64, requires_grad=True)
    packed = MRFSimulator().packed(flip=flip, TR=tr)
    events = {"duration": packed.duration.double(), "flip": packed.flip.double(), "phase": packed.phase.detach(),
              "kind": packed.kind, "action": packed.action, "output_index": packed.output_index}
    base = tuple(torch.tensor(v, dtype=torch.float64) for v in (t1, t2, np.ones(atoms)))

    def run(a, b, c):
        return P.simulate(events, {"t1": a, "t2": b, "m0": c}, states=states, outputs=pulses)

    cols = []
    for k in range(3):
        tangents = tuple(torch.ones(atoms, dtype=torch.float64) if j == k else torch.zeros(atoms, dtype=torch.float64)
                         for j in range(3))
        cols.append(1j * torch.func.jvp(run, base, tangents)[1])
    jac = torch.stack(cols, 1)
    rows = torch.cat((jac.real, jac.imag), -1)
    loss = torch.diagonal(torch.linalg.inv(rows @ rows.transpose(-1, -2)), dim1=-2, dim2=-1).sum()
    loss.backward()
    return float(loss.detach()), flip.grad.numpy(), tr.grad.numpy()


def test_reference_design_gradient_matches_fp64_autograd():
    """The reference's FP32 forward-over-reverse passes, driven as its
    autograd would drive them, agree with FP64 autograd of the oracle to
    FP32 rounding (measured ~3e-7)."""
    _reference_or_skip()
    flips, trs, t1, t2 = design_inputs(60, 3)
    loss, g_flip, g_tr = reference_gradient(flips, trs, t1, t2)
    ref_loss, ref_flip, ref_tr = _fp64_design_gradient(flips, trs, t1, t2)
    assert abs(loss - ref_loss) / ref_loss < 1e-5
    assert _cases.rel_err(g_flip, ref_flip) < 1e-5
    assert _cases.rel_err(g_tr, ref_tr) < 1e-5
    assert g_tr[-1] == 0.0


def test_reference_vjp_is_the_transpose_of_the_reference_jvp():
    """<J v, w> = <v, J^T w> between the reference's own forward and reverse
    passes, along T1 and along the event flips (complex path, phases on)."""
    _reference_or_skip()
    from torchsim_amd.simulators import MRFSimulator

    events = oracle.events_of(MRFSimulator().packed(flip=torch.tensor(_cases.benchmark_flips(40)), TR=10.0))
    rng = np.random.default_rng(7)
    events["phase"] = np.where(events["kind"] == 1, rng.uniform(-2, 2, len(events["kind"])), 0.0).astype(np.float32)
    tissue = {"t1": np.array([900.0, 1300.0]), "t2": np.array([60.0, 110.0]), "b0": np.array([3.0, -2.0])}
    w = rng.normal(size=(2, 40)) + 1j * rng.normal(size=(2, 40))
    grads = oracle.reference_vjp(events, tissue, w, states=10, outputs=40, real_axis=-1)
    for direction, read in (({"t1": np.array([1.0, 0.5])}, lambda g: g["t1"] @ np.array([1.0, 0.5])),
                            ({"flip": (v := rng.normal(size=len(events["kind"])))}, lambda g: g["flip"] @ v)):
        jv = oracle.reference_complex(events, tissue, states=10, outputs=40, tangent=direction)
        lhs = float((jv.conj() * w).real.sum())
        assert abs(lhs - read(grads)) / abs(lhs) < 1e-4


def test_design_gradient_fixture_is_the_reference_gradient():
    """The committed fixture is what the reference computes today (so the GPU
    test may read it on a box without oracle/_ref)."""
    _reference_or_skip()
    golden = np.load(GOLDEN)
    flips, trs, t1, t2 = design_inputs(880, 1024)
    loss, g_flip, g_tr = reference_gradient(flips, trs, t1, t2)
    assert abs(loss - float(golden["loss"])) / loss < 1e-12
    assert np.array_equal(g_flip, golden["grad_flip"])
    assert np.array_equal(g_tr, golden["grad_tr"])
