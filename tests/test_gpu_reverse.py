"""Reverse mode through the real-subspace adjoint kernel with respect to the
tissue (T1, T2, M0, B1, inversion efficiency) as well as the sequence: a
signal loss (first order) and the Cramer-Rao loss of the Jacobian
(forward-over-reverse), against FP64 autograd of the torch oracle and, where
oracle/_ref is built, against the reference's own simulate_vjp /
simulate_vjp_jvp (reference sequence/_accelerators.py:1544-1586,
src/torchsim/_epg_cpu.cpp:11150-11755)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import packed as P
from tests import _cases

pytestmark = pytest.mark.gpu
TOL = 1e-5  # max|g - g_ref| / max|g_ref|; a dropped term or a wrong sign is off by far more


def _inputs(pulses, atoms, seed):
    rng = np.random.default_rng(seed)
    flips = rng.uniform(2.0, 70.0, pulses)
    trs = rng.uniform(8.0, 15.0, pulses)
    t1, t2 = _cases.random_tissue(atoms, seed=seed)
    m0 = rng.uniform(0.5, 1.0, atoms).astype(np.float32)
    b1 = rng.uniform(0.7, 1.3, atoms).astype(np.float32)
    ie = rng.uniform(0.8, 1.0, atoms).astype(np.float32)
    return flips, trs, {"T1": t1, "T2": t2, "M0": m0, "B1": b1, "inv_efficiency": ie}


_ORACLE_NAME = {"T1": "t1", "T2": "t2", "M0": "m0", "B1": "b1", "inv_efficiency": "ie"}


def _fp64_events(flip, tr):
    from torchsim_amd.simulators import MRFSimulator

    packed = MRFSimulator().packed(flip=flip, TR=tr)
    return {"duration": packed.duration.double(), "flip": packed.flip.double(), "phase": packed.phase.detach(),
            "kind": packed.kind, "action": packed.action, "output_index": packed.output_index}


def test_signal_loss_reaches_every_tissue_property_and_the_sequence():
    from torchsim_amd.simulators import MRFSimulator

    flips, trs, tissue = _inputs(120, 24, seed=41)
    rng = np.random.default_rng(42)
    weight = rng.normal(size=(24, 120))
    flip = torch.tensor(flips, dtype=torch.float32, requires_grad=True)
    tr = torch.tensor(trs, dtype=torch.float32, requires_grad=True)
    maps = {k: torch.tensor(v, device="cuda", requires_grad=True) for k, v in tissue.items()}
    sig = MRFSimulator(states=12).simulate(flip=flip, TR=tr, **maps)
    loss = (sig.real * torch.tensor(weight, dtype=torch.float32, device="cuda")).sum()
    loss.backward()

    flip64 = torch.tensor(flips, dtype=torch.float64, requires_grad=True)
    tr64 = torch.tensor(trs, dtype=torch.float64, requires_grad=True)
    maps64 = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in tissue.items()}
    ref = 1j * P.simulate(_fp64_events(flip64, tr64), {_ORACLE_NAME[k]: v for k, v in maps64.items()}, states=12,
                          outputs=120)
    (ref.real * torch.tensor(weight)).sum().backward()
    errors = {k: _cases.rel_err(maps[k].grad.cpu().numpy(), maps64[k].grad.numpy()) for k in tissue}
    errors["flip"] = _cases.rel_err(flip.grad.numpy(), flip64.grad.numpy())
    errors["TR"] = _cases.rel_err(tr.grad.numpy(), tr64.grad.numpy())
    print("signal-loss gradients vs FP64 autograd:", {k: f"{v:.2e}" for k, v in errors.items()})
    assert max(errors.values()) < TOL, errors


def test_signal_loss_tissue_gradients_match_the_reference_vjp():
    """The same per-atom gradients from the reference's own first-order pass."""
    if oracle.reference_module() is None:
        pytest.skip("oracle/_ref is not built")
    from torchsim_amd.simulators import MRFSimulator

    flips, trs, tissue = _inputs(200, 40, seed=43)
    weight = np.random.default_rng(44).normal(size=(40, 200))
    maps = {k: torch.tensor(v, device="cuda", requires_grad=True) for k, v in tissue.items()}
    seq = MRFSimulator(flip=torch.tensor(flips, dtype=torch.float32), TR=torch.tensor(trs, dtype=torch.float32),
                       states=10)
    sig = seq.simulate(**maps)
    (sig.real * torch.tensor(weight, dtype=torch.float32, device="cuda")).sum().backward()
    events = oracle.events_of(seq.packed(**seq.bound))
    ref_tissue = {_ORACLE_NAME[k]: v for k, v in tissue.items()}
    # public = 1j * (i m0 a0) = -(imaginary plane): the plane's cotangent is -weight
    grads = oracle.reference_vjp(events, ref_tissue, -1j * weight, states=10, outputs=200)
    for k in tissue:
        err = _cases.rel_err(maps[k].grad.cpu().numpy(), grads[_ORACLE_NAME[k]])
        print(f"{k}: {err:.2e}")
        assert err < TOL, k


def _crlb_gpu(flips, trs, tissue, names, diff, states=10):
    from torchsim_amd.optim import crlb
    from torchsim_amd.simulators import MRFSimulator

    flip = torch.tensor(flips, dtype=torch.float32, requires_grad=True)
    tr = torch.tensor(trs, dtype=torch.float32, requires_grad=True)
    maps = {k: torch.tensor(v, device="cuda", requires_grad=k in diff) for k, v in tissue.items()}
    _, jac = MRFSimulator(states=states).jacobian(names, flip=flip, TR=tr, **maps)
    loss = crlb(jac).sum()
    loss.backward()
    out = {k: maps[k].grad.cpu().numpy().astype(np.float64) for k in diff}
    out["flip"] = flip.grad.numpy().astype(np.float64)
    out["TR"] = tr.grad.numpy().astype(np.float64)
    return float(loss), out


def _crlb_fp64(flips, trs, tissue, names, diff, states=10):
    flip = torch.tensor(flips, dtype=torch.float64, requires_grad=True)
    tr = torch.tensor(trs, dtype=torch.float64, requires_grad=True)
    maps = {k: torch.tensor(v, dtype=torch.float64, requires_grad=k in diff) for k, v in tissue.items()}
    events = _fp64_events(flip, tr)
    atoms = len(tissue["T1"])
    primal = tuple(maps[k] for k in names)

    def run(*values):
        given = {_ORACLE_NAME[k]: v for k, v in maps.items()}
        given.update({_ORACLE_NAME[k]: v for k, v in zip(names, values)})
        return P.simulate(events, given, states=states, outputs=len(flips))

    cols = []
    for j in range(len(names)):
        tangents = tuple(torch.ones(atoms, dtype=torch.float64) if q == j else torch.zeros(atoms, dtype=torch.float64)
                         for q in range(len(names)))
        cols.append(1j * torch.func.jvp(run, primal, tangents)[1])
    jac = torch.stack(cols, 1)
    rows = torch.cat((jac.real, jac.imag), -1)
    loss = torch.diagonal(torch.linalg.inv(rows @ rows.transpose(-1, -2)), dim1=-2, dim2=-1).sum()
    loss.backward()
    out = {k: maps[k].grad.numpy() for k in diff}
    out["flip"] = flip.grad.numpy()
    out["TR"] = tr.grad.numpy()
    return float(loss.detach()), out


@pytest.mark.parametrize("names", [("T1", "T2", "M0"), ("T1", "T2", "B1", "M0")])
def test_crlb_gradient_reaches_the_tissue(names):
    """Forward-over-reverse: the Cramer-Rao loss of the Jacobian along
    ``names`` differentiated with respect to T1, T2, B1 and the sequence,
    against FP64 autograd through forward-mode of the oracle (the B1 column
    exercises the rotation's second derivative)."""
    flips, trs, tissue = _inputs(60, 4, seed=45)
    diff = ("T1", "T2", "B1")
    loss, ours = _crlb_gpu(flips, trs, tissue, names, diff)
    ref_loss, ref = _crlb_fp64(flips, trs, tissue, names, diff)
    errors = {k: _cases.rel_err(ours[k], ref[k]) for k in ref}
    print(f"crlb{names} gradients vs FP64 autograd:", {k: f"{v:.2e}" for k, v in errors.items()})
    assert abs(loss - ref_loss) / ref_loss < 1e-4
    assert max(errors.values()) < TOL, errors


def test_crlb_tissue_gradients_match_the_reference_vjp_jvp():
    """T1 and T2 gradients of the design loss against the reference's
    forward-over-reverse passes (one per tissue column plus the first-order
    pass for the M0 column)."""
    if oracle.reference_module() is None:
        pytest.skip("oracle/_ref is not built")
    from torchsim_amd.simulators import MRFSimulator

    flips, trs, tissue = _inputs(300, 64, seed=47)
    tissue = {k: tissue[k] for k in ("T1", "T2")}
    tissue["M0"] = np.ones(64, dtype=np.float32)
    loss, ours = _crlb_gpu(flips, trs, tissue, ("T1", "T2", "M0"), ("T1", "T2"))
    packed = MRFSimulator().packed(flip=torch.tensor(flips, dtype=torch.float32),
                                   TR=torch.tensor(trs, dtype=torch.float32))
    events = oracle.events_of(packed)
    # the cotangents on the columns, from the reference's own columns (as in
    # oracle.reference_crlb_gradient), then the tissue rows of each pass
    t1, t2 = tissue["T1"], tissue["T2"]
    base = {"t1": t1, "t2": t2, "m0": np.ones_like(t1)}
    kw = dict(states=10, outputs=300, features=oracle.FEATURE_ALL)
    planes = [oracle.reference_real(events, base, tangent={n: 1.0}, **kw) for n in ("t1", "t2")]
    planes.append(oracle.reference_real(events, base, **kw))
    jac = -np.stack(planes, 1).astype(np.float64)
    fisher = jac @ np.transpose(jac, (0, 2, 1))
    inverse = np.linalg.inv(fisher)
    d_plane = -(-2.0 * (inverse @ inverse) @ jac)
    g = {"t1": np.zeros(64), "t2": np.zeros(64)}
    for k, name in enumerate(("t1", "t2")):
        r = oracle.reference_vjp(events, base, 1j * d_plane[:, k], tangent={name: 1.0}, **kw)
        g["t1"] += r["t1"]
        g["t2"] += r["t2"]
    r = oracle.reference_vjp(events, base, 1j * d_plane[:, 2], **kw)
    g["t1"] += r["t1"]
    g["t2"] += r["t2"]
    for name, key in (("t1", "T1"), ("t2", "T2")):
        err = _cases.rel_err(ours[key], g[name])
        print(f"{key}: {err:.2e}")
        assert err < TOL, key


def test_what_the_complex_adjoint_does_not_carry_raises():
    """The complex path's forward-over-reverse passes gradients to the
    sequence only, and its adjoint reaches no pool-b property: a Jacobian
    under reverse mode with a tissue map that requires grad, or a pool-b map
    that requires grad, is refused rather than silently cut off."""
    from torchsim_amd.simulators import MRFSimulator

    t1 = torch.tensor([800.0, 1200.0], device="cuda", requires_grad=True)
    t2 = torch.tensor([60.0, 80.0], device="cuda")
    seq = MRFSimulator(flip=torch.full((20,), 30.0), TR=10.0, phases=torch.linspace(0, 90, 20))
    with pytest.raises(NotImplementedError, match="tissue maps are constants"):
        seq.jacobian(("T1", "T2"), T1=t1, T2=t2)
    fraction = torch.tensor([0.1, 0.2], device="cuda", requires_grad=True)
    with pytest.raises(NotImplementedError, match="pool_b_fraction"):
        seq.simulate(T1=t1, T2=t2, poolB_fraction=fraction, poolB_exchange=10.0, poolB_T1=400.0, poolB_T2=15.0)


@pytest.mark.parametrize("protocol", ["mrf", "fse", "inverted_balanced"])
def test_host_op_count_sizes_the_adjoint_exactly(protocol):
    """The adjoint's workspace is sized by the op count worked out on the
    host; the gradients equal, bitwise, those of a call that reads the count
    back from the device (a host count that fell short would drop ops)."""
    from torchsim_amd.sequence import _engine
    from torchsim_amd.simulators import FSESimulator, MRFSimulator, TrueFISPSimulator

    if protocol == "mrf":
        packed = MRFSimulator().packed(flip=torch.tensor(_cases.benchmark_flips(90)), TR=10.0)
    elif protocol == "fse":
        packed = FSESimulator().packed(flip=torch.linspace(180, 120, 24), ESP=5.0, TR=3000.0)
    else:
        packed = TrueFISPSimulator().packed(flip=40.0, TR=5.0, nshots=50, inversion="ideal", spoiler=4.0)
    t1, t2 = _cases.random_tissue(17, seed=48)
    tissue = _engine.prepare_tissue({"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2)}, torch.device("cuda", 0),
                                    ("t1_ms", "t2_ms"))
    outputs = packed.output_count
    weight = torch.tensor(np.random.default_rng(49).normal(size=(17, outputs)), dtype=torch.float32, device="cuda")
    runs = []
    for bound in (None, -1):
        runs.append(_engine.adjoint_real(packed, tissue, states=8, columns=(), flip=packed.flip,
                                         duration=packed.duration, signal_cotangent=weight, jacobian_cotangent=None,
                                         tissue_wanted=("t1_ms", "t2_ms"), op_bound=bound))
    (f0, d0, g0), (f1, d1, g1) = runs
    assert _engine.op_count(packed) > 0
    assert torch.equal(f0, f1) and torch.equal(d0, d1)
    assert torch.equal(g0["t1_ms"], g1["t1_ms"]) and torch.equal(g0["t2_ms"], g1["t2_ms"])
    assert float(g0["t1_ms"].abs().max()) > 0


def test_an_inversion_time_of_zero_has_its_gradient():
    """TI = 0 (the MRF default) leaves the first excitation's interval empty;
    its duration gradient -- the pairing of the adjoint with the relaxation
    generator there -- still reaches TI, as FP64 autograd says."""
    from torchsim_amd.simulators import MRFSimulator

    flips, trs, tissue = _inputs(60, 12, seed=51)
    weight = np.random.default_rng(52).normal(size=(12, 60))
    ti = torch.tensor(0.0, requires_grad=True)
    maps = {k: torch.tensor(v, device="cuda") for k, v in tissue.items()}
    maps["T1"].requires_grad_()
    sig = MRFSimulator(states=10).simulate(flip=torch.tensor(flips, dtype=torch.float32),
                                           TR=torch.tensor(trs, dtype=torch.float32), TI=ti, **maps)
    (sig.real * torch.tensor(weight, dtype=torch.float32, device="cuda")).sum().backward()

    ti64 = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    packed = MRFSimulator().packed(flip=torch.tensor(flips), TR=torch.tensor(trs), TI=ti64)
    events = {"duration": packed.duration.double(), "flip": packed.flip.double(), "phase": packed.phase.detach(),
              "kind": packed.kind, "action": packed.action, "output_index": packed.output_index}
    ref = 1j * P.simulate(events, {_ORACLE_NAME[k]: torch.tensor(v, dtype=torch.float64) for k, v in tissue.items()},
                          states=10, outputs=60)
    (ref.real * torch.tensor(weight)).sum().backward()
    assert float(ti64.grad) != 0.0
    err = abs(float(ti.grad) - float(ti64.grad)) / abs(float(ti64.grad))
    print(f"d/dTI: {float(ti.grad):.6e} vs FP64 {float(ti64.grad):.6e} ({err:.1e})")
    assert err < TOL


def test_every_duration_gradient_matches_the_reference_vjp():
    """With every zero interval asked for (the reference's convention), each
    event's duration gradient -- the sampled and crusher events' included --
    is the reference's simulate_vjp's, first order and forward-over-reverse."""
    if oracle.reference_module() is None:
        pytest.skip("oracle/_ref is not built")
    from torchsim_amd.sequence import _engine
    from torchsim_amd.simulators import MRFSimulator

    flips, trs, tissue = _inputs(80, 16, seed=53)
    packed = MRFSimulator().packed(flip=torch.tensor(flips, dtype=torch.float32),
                                   TR=torch.tensor(trs, dtype=torch.float32))
    events = oracle.events_of(packed)
    t1, t2 = tissue["T1"], tissue["T2"]
    prepared = _engine.prepare_tissue({"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2)},
                                      torch.device("cuda", 0), ("t1_ms", "t2_ms"))
    rng = np.random.default_rng(54)
    ws = rng.normal(size=(16, 80))
    wj = rng.normal(size=(16, 1, 80))
    # first order
    _, gd, _ = _engine.adjoint_real(packed, prepared, states=10, columns=(), flip=packed.flip,
                                    duration=packed.duration, signal_cotangent=torch.tensor(ws),
                                    jacobian_cotangent=None, zero_intervals=True)
    ref = oracle.reference_vjp(events, {"t1": t1, "t2": t2}, 1j * ws, states=10, outputs=80)
    assert np.count_nonzero(ref["duration"][np.asarray(events["duration"]) == 0]) > 0
    err = _cases.rel_err(gd.cpu().numpy(), ref["duration"])
    print(f"first-order duration gradient vs reference: {err:.2e}")
    assert err < 1e-4
    # forward-over-reverse along T1
    _, gd, _ = _engine.adjoint_real(packed, prepared, states=10, columns=("t1_ms",), flip=packed.flip,
                                    duration=packed.duration, signal_cotangent=None,
                                    jacobian_cotangent=torch.tensor(wj), zero_intervals=True)
    ref = oracle.reference_vjp(events, {"t1": t1, "t2": t2}, 1j * wj[:, 0], states=10, outputs=80,
                               tangent={"t1": 1.0})
    err = _cases.rel_err(gd.cpu().numpy(), ref["duration"])
    print(f"forward-over-reverse duration gradient vs reference: {err:.2e}")
    assert err < 1e-4
