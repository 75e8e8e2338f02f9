"""Forward-over-reverse on the complex kernel: the gradient, with respect to
the sequence's flips, TRs and phases, of a loss that reads the Jacobian of a
phase-cycled, off-resonant, exchanging or diffusing protocol -- against the
reference's own FP32 ``simulate_vjp_jvp`` (real_axis -1, built from source)
per tangent, and against FP64 autograd through forward mode of the oracle
for the whole loss (signal and columns T1, T2, M0, B1) and for the
Cramer-Rao design loss."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import packed as P
from tests import _cases

pytestmark = pytest.mark.gpu

TOL = 2e-5  # relative to the largest gradient entry, vs FP64


def _protocol(pulses, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(5.0, 60.0, pulses), rng.uniform(8.0, 15.0, pulses), rng.uniform(0.0, 360.0, pulses)


def _rel(x, ref):
    return float(np.abs(np.asarray(x) - ref).max() / np.abs(ref).max())


@pytest.mark.parametrize("exchanging", [False, True])
def test_each_tangent_matches_the_reference_simulate_vjp_jvp(exchanging):
    if oracle.reference_module() is None:
        pytest.skip("oracle/_ref is not built")
    from torchsim_amd.sequence import _complex, _engine
    from torchsim_amd.simulators import MRFSimulator

    atoms, pulses = 16, 80
    t1, t2 = _cases.random_tissue(atoms, seed=111)
    rng = np.random.default_rng(112)
    flips, trs, phases = _protocol(pulses, 113)
    b0 = rng.uniform(-25.0, 25.0, atoms).astype(np.float32)
    b1 = rng.uniform(0.8, 1.2, atoms).astype(np.float32)
    weight = rng.normal(size=(atoms, pulses)) + 1j * rng.normal(size=(atoms, pulses))
    packed = MRFSimulator().packed(flip=torch.tensor(flips, dtype=torch.float32),
                                   TR=torch.tensor(trs, dtype=torch.float32),
                                   phases=torch.tensor(phases, dtype=torch.float32))
    events = oracle.events_of(packed)
    fields = {"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2), "b0_hz": torch.tensor(b0),
              "b1": torch.tensor(b1)}
    ref_tissue = {"t1": t1, "t2": t2, "b0": b0, "b1": b1}
    if exchanging:
        fraction = rng.uniform(0.05, 0.3, atoms).astype(np.float32)
        exchange = rng.uniform(5.0, 20.0, atoms).astype(np.float32)
        fields.update(pool_b_fraction=torch.tensor(fraction), pool_b_exchange_rate=torch.tensor(exchange),
                      t1_pool_b_ms=400.0, t2_pool_b_ms=15.0, pool_b_shift_hz=35.0)
        ref_tissue.update(pool_b_fraction=fraction, pool_b_exchange=exchange, t1_pool_b=400.0, t2_pool_b=15.0,
                          pool_b_shift=35.0)
    names = ("t1_ms", "t2_ms", "b1")
    tissue = _engine.prepare_tissue(fields, torch.device("cuda", 0), names)
    for name, ref_name in zip(names, ("t1", "t2", "b1")):
        cot = torch.zeros((atoms, 1, pulses), dtype=torch.complex64)
        cot[:, 0] = torch.tensor(weight, dtype=torch.complex64)
        gf, gp, gd = _complex.adjoint_complex_jacobian(
            packed, tissue, states=10, factor=1j, columns=(name,), signal_cotangent=None, jacobian_cotangent=cot)
        ref = oracle.reference_vjp(events, ref_tissue, np.conj(1j) * weight, states=10, outputs=pulses,
                                   real_axis=-1, exchanging=exchanging, tangent={ref_name: 1.0})
        errors = {"flip": _rel(gf.cpu().numpy(), ref["flip"]), "phase": _rel(gp.cpu().numpy(), ref["phase"]),
                  "duration": _rel(gd.cpu().numpy(), ref["duration"])}
        print(f"exchanging={exchanging} along {name}:", {k: f"{v:.2e}" for k, v in errors.items()})
        assert max(errors.values()) < 1e-4, (name, errors)


def _loss_both(pulses, atoms, seed, names, *, fixed=None, oracle_fixed=None, run=None, crlb=False):
    """The loss on the GPU (FP32, through Simulator.jacobian and its
    forward-over-reverse backward) and in FP64 on the oracle, with their
    gradients along the flips, TRs and phases."""
    from torchsim_amd.optim import crlb as crlb_loss
    from torchsim_amd.simulators import MRFSimulator

    fixed, oracle_fixed, run = fixed or {}, oracle_fixed or {}, run or {}
    flips, trs, phases = _protocol(pulses, seed)
    rng = np.random.default_rng(seed + 1)
    t1, t2 = _cases.random_tissue(atoms, seed=seed)
    tissue = {"T1": t1, "T2": t2, "M0": rng.uniform(0.5, 1.0, atoms).astype(np.float32),
              "B1": rng.uniform(0.8, 1.2, atoms).astype(np.float32),
              "B0": rng.uniform(-20.0, 20.0, atoms).astype(np.float32)}
    ws = rng.normal(size=(atoms, pulses)) + 1j * rng.normal(size=(atoms, pulses))
    wj = rng.normal(size=(atoms, len(names), pulses)) + 1j * rng.normal(size=(atoms, len(names), pulses))

    flip = torch.tensor(flips, dtype=torch.float32, requires_grad=True)
    tr = torch.tensor(trs, dtype=torch.float32, requires_grad=True)
    ph = torch.tensor(phases, dtype=torch.float32, requires_grad=True)
    maps = {k: torch.tensor(v, device="cuda") for k, v in tissue.items()}
    signal, jac = MRFSimulator(states=10).jacobian(names, flip=flip, TR=tr, phases=ph, **maps, **fixed, **run)
    if crlb:
        loss = crlb_loss(jac).sum()
    else:
        loss = ((signal * torch.tensor(ws, dtype=torch.complex64, device="cuda").conj()).real.sum()
                + (jac * torch.tensor(wj, dtype=torch.complex64, device="cuda").conj()).real.sum())
    loss.backward()
    gpu = {"flip": flip.grad.numpy(), "TR": tr.grad.numpy(), "phases": ph.grad.numpy()}

    flip64 = torch.tensor(flips, dtype=torch.float64, requires_grad=True)
    tr64 = torch.tensor(trs, dtype=torch.float64, requires_grad=True)
    ph64 = torch.tensor(phases, dtype=torch.float64, requires_grad=True)
    packed = MRFSimulator().packed(flip=flip64, TR=tr64, phases=ph64)
    events = {"duration": packed.duration.double(), "flip": packed.flip.double(), "phase": packed.phase.double(),
              "kind": packed.kind, "action": packed.action, "output_index": packed.output_index}
    oracle_name = {"T1": "t1", "T2": "t2", "M0": "m0", "B1": "b1", "B0": "b0"}
    base = {oracle_name[k]: torch.tensor(v, dtype=torch.float64) for k, v in tissue.items()}
    base.update({k: (torch.tensor(v, dtype=torch.float64) if isinstance(v, np.ndarray) else v)
                 for k, v in oracle_fixed.items()})
    primal = tuple(base[oracle_name[k]] for k in names)

    def simulate(*values):
        given = dict(base)
        given.update({oracle_name[k]: v for k, v in zip(names, values)})
        return 1j * P.simulate(events, given, states=10, outputs=pulses)

    sig64 = simulate(*primal)
    cols = []
    for j in range(len(names)):
        tangents = tuple(torch.ones(atoms, dtype=torch.float64) if q == j else torch.zeros(atoms, dtype=torch.float64)
                         for q in range(len(names)))
        cols.append(torch.func.jvp(simulate, primal, tangents)[1])
    jac64 = torch.stack(cols, 1)
    if crlb:
        rows = torch.cat((jac64.real, jac64.imag), -1)
        ref_loss = torch.diagonal(torch.linalg.inv(rows @ rows.transpose(-1, -2)), dim1=-2, dim2=-1).sum()
    else:
        ref_loss = (sig64 * torch.tensor(ws).conj()).real.sum() + (jac64 * torch.tensor(wj).conj()).real.sum()
    ref_loss.backward()
    ref = {"flip": flip64.grad.numpy(), "TR": tr64.grad.numpy(), "phases": ph64.grad.numpy()}
    return float(loss), float(ref_loss.detach()), gpu, ref


def test_signal_and_jacobian_loss_against_fp64_autograd():
    """Phase cycling, off-resonance and a B1 map: columns T1, T2, M0, B1."""
    loss, ref_loss, gpu, ref = _loss_both(70, 12, 121, ("T1", "T2", "M0", "B1"))
    errors = {k: _rel(gpu[k], ref[k]) for k in ref}
    print("phase-cycled signal + Jacobian loss:", {k: f"{v:.2e}" for k, v in errors.items()})
    assert max(errors.values()) < TOL, errors


def test_exchanging_pool_jacobian_loss_against_fp64_autograd():
    fixed = {"poolB_fraction": torch.full((12,), 0.15, device="cuda"),
             "poolB_exchange": torch.full((12,), 12.0, device="cuda"), "poolB_T1": 400.0, "poolB_T2": 15.0,
             "poolB_shift": 35.0}
    oracle_fixed = {"fraction": np.full(12, 0.15), "exchange": np.full(12, 12.0), "t1b": 400.0, "t2b": 15.0,
                    "shift": 35.0}
    loss, ref_loss, gpu, ref = _loss_both(60, 12, 131, ("T1", "T2", "M0"), fixed=fixed, oracle_fixed=oracle_fixed)
    errors = {k: _rel(gpu[k], ref[k]) for k in ref}
    print("exchanging signal + Jacobian loss:", {k: f"{v:.2e}" for k, v in errors.items()})
    assert max(errors.values()) < TOL, errors


def test_diffusing_crlb_design_loss_against_fp64_autograd():
    """The Cramer-Rao design loss of (T1, T2, M0) on a phase-cycled,
    off-resonant protocol with diffusion through a crusher geometry."""
    from torchsim_amd.model._simulator import damping_scale

    run = {"crusher_dephasing_rad": 8 * np.pi, "voxel_size_m": 0.5e-3}
    fixed = {"D": torch.linspace(0.5, 3.0, 10, device="cuda")}
    oracle_fixed = {"diffusion": np.linspace(0.5, 3.0, 10) * damping_scale(run)}
    loss, ref_loss, gpu, ref = _loss_both(60, 10, 141, ("T1", "T2", "M0"), fixed=fixed, oracle_fixed=oracle_fixed,
                                          run=run, crlb=True)
    errors = {k: _rel(gpu[k], ref[k]) for k in ref}
    print(f"diffusing crlb loss {loss:.6e} vs {ref_loss:.6e}:", {k: f"{v:.2e}" for k, v in errors.items()})
    assert abs(loss - ref_loss) / ref_loss < 1e-4
    assert max(errors.values()) < TOL, errors


def test_checkpoint_length_changes_nothing():
    from torchsim_amd.sequence import _complex, _engine
    from torchsim_amd.simulators import MRFSimulator

    atoms, pulses = 10, 50
    t1, t2 = _cases.random_tissue(atoms, seed=151)
    flips, trs, phases = _protocol(pulses, 152)
    packed = MRFSimulator().packed(flip=torch.tensor(flips, dtype=torch.float32),
                                   TR=torch.tensor(trs, dtype=torch.float32),
                                   phases=torch.tensor(phases, dtype=torch.float32))
    tissue = _engine.prepare_tissue({"t1_ms": torch.tensor(t1), "t2_ms": torch.tensor(t2)},
                                    torch.device("cuda", 0), ("t1_ms", "t2_ms"))
    rng = np.random.default_rng(153)
    ws = torch.tensor(rng.normal(size=(atoms, pulses)), dtype=torch.complex64)
    wj = torch.tensor(rng.normal(size=(atoms, 3, pulses)), dtype=torch.complex64)
    runs = [_complex.adjoint_complex_jacobian(packed, tissue, states=8, factor=1j, columns=("t1_ms", "t2_ms", "m0"),
                                              signal_cotangent=ws, jacobian_cotangent=wj, segment_events=k)
            for k in (0, 1, 7, 10000)]
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert torch.equal(a, b)
