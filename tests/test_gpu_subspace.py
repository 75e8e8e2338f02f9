"""Projection onto a temporal basis fused into the real-subspace kernel,
against the FP64 oracle's signal and Jacobian projected in FP64 (signal
coefficients within 1e-5, Jacobian coefficients within 1e-4 of the tensor's
largest magnitude), and against the full simulation projected on the device."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import subspace as osub
from tests import _cases
from tests.test_gpu_real import _oracle_mrf

pytestmark = pytest.mark.gpu

SIGNAL_TOL = 1e-5
JACOBIAN_TOL = 1e-4


def _basis(signals: np.ndarray, rank: int) -> np.ndarray:
    basis, _ = osub.fit(signals, rank)
    return basis


@pytest.mark.parametrize("rank", [1, 5, 8, 16])
def test_mrf_dictionary_coefficients_match_the_oracle(rank):
    """Config-1 physics (S = 20, B1 and M0 maps), coefficients of the signal
    and of d/d(T1, T2, B1, M0) for a ragged atom count."""
    from torchsim_amd import Subspace
    from torchsim_amd.simulators import MRFSimulator

    print(f"case rank={rank}", flush=True)
    flip = _cases.mrf_flips()
    T1, T2, B1 = _cases.dictionary_grid()
    rng = np.random.default_rng(rank)
    pick = rng.choice(T1.size, 1203, replace=False)
    t1, t2, b1 = T1[pick], T2[pick], B1[pick]
    m0 = rng.uniform(0.5, 1.5, pick.size)
    names = ("T1", "T2", "B1", "M0")
    ref_sig, ref_jac = _oracle_mrf(flip, 10.0, t1, t2, b1=b1, m0=m0, states=20, names=names)
    basis = _basis(ref_sig[:400].astype(np.complex128), rank)
    subspace = Subspace(basis=torch.from_numpy(basis.astype(np.complex64)),
                        singular_values=torch.ones(rank))
    seq = MRFSimulator(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0, states=20)
    on = {name: torch.tensor(v, dtype=torch.float32).cuda() for name, v in
          (("T1", t1), ("T2", t2), ("B1", b1), ("M0", m0))}
    coef, jcoef = seq.project(subspace, names, **on)
    assert coef.shape == (1203, rank) and jcoef.shape == (1203, 4, rank)
    assert coef.dtype == torch.complex64
    b64 = subspace.basis.numpy().astype(np.complex128)
    assert _cases.rel_err(coef.cpu().numpy(), osub.project(ref_sig, b64)) < SIGNAL_TOL
    for p in range(4):
        assert _cases.rel_err(jcoef[:, p].cpu().numpy(), osub.project(ref_jac[:, p], b64)) < JACOBIAN_TOL
    # the signal alone
    only = seq.project(subspace, **on)
    torch.testing.assert_close(only, coef, rtol=0, atol=0)


def _unfused(sim, subspace, names, values):
    signal, jacobian = sim.jacobian(names, **values)
    return subspace.project(signal), subspace.project(jacobian)


@pytest.mark.parametrize("states", [3, 10, 20, 40])
def test_bssfp_projection_equals_the_projected_simulation(states):
    """A balanced train (pulses alternating in phase, so the kernel's sample
    signs are in play) at several state counts, so several tiles run."""
    from torchsim_amd import Subspace
    from torchsim_amd.simulators import TrueFISPSimulator

    print(f"case states={states}", flush=True)
    t1, t2 = _cases.random_tissue(777, seed=states)
    sim = TrueFISPSimulator(flip=50.0, TR=4.5, nshots=300, states=states)
    values = {"T1": torch.tensor(t1).cuda(), "T2": torch.tensor(t2).cuda()}
    subspace = Subspace.fit(sim.simulate(**values).cpu(), 6)
    coef, jcoef = sim.project(subspace, ("T1", "T2"), **values)
    want, jwant = _unfused(sim, subspace, ("T1", "T2"), values)
    assert _cases.rel_err(coef.cpu().numpy(), want.cpu().numpy()) < SIGNAL_TOL
    for p in range(2):
        assert _cases.rel_err(jcoef[:, p].cpu().numpy(), jwant[:, p].cpu().numpy()) < JACOBIAN_TOL


@pytest.mark.parametrize("tile,rank", [(1, 4), (5, 8), (10, 16), (10, 3)])
def test_every_lane_width_through_the_c_abi(tile, rank):
    """Explicit tiles of 1, 2 and 4 lanes per atom (4 directions per lane),
    S = 8; a rank beyond a tile's lanes, or beyond 16, is refused."""
    from torchsim_amd.sequence import _engine
    from torchsim_amd.simulators import MRFSimulator

    print(f"case tile={tile} rank={rank}", flush=True)
    flip = _cases.mrf_flips()[:200]
    t1, t2 = _cases.random_tissue(513, seed=tile)
    packed = MRFSimulator().packed(flip=torch.tensor(flip, dtype=torch.float32), TR=10.0)
    fields = {"t1_ms": torch.tensor(t1, dtype=torch.float32).cuda(),
              "t2_ms": torch.tensor(t2, dtype=torch.float32).cuda()}
    tissue = _engine.prepare_tissue(fields, torch.device("cuda"), ("t1_ms", "t2_ms"))
    signal, jacobian = _engine.simulate_real(packed, tissue, states=8, columns=("t1_ms", "t2_ms"), factor=1j,
                                             tile=tile)
    basis = torch.from_numpy(_basis(signal[:200].cpu().numpy().astype(np.complex128), rank)
                             .astype(np.complex64)).cuda()
    coef, jcoef = _engine.simulate_projected(packed, tissue, basis, states=8, columns=("t1_ms", "t2_ms"),
                                             factor=1j, tile=tile)
    assert _cases.rel_err(coef.cpu().numpy(), (signal @ basis).cpu().numpy()) < SIGNAL_TOL
    for p in range(2):
        assert _cases.rel_err(jcoef[:, p].cpu().numpy(), (jacobian[:, p] @ basis).cpu().numpy()) < JACOBIAN_TOL
    lanes = {1: 1, 5: 2, 10: 4}[tile]
    too_many = torch.zeros(packed.output_count, 4 * lanes + 1, dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError if lanes < 4 else NotImplementedError, match="lanes per atom|between 1 and 16"):
        _engine.simulate_projected(packed, tissue, too_many, states=8, factor=1j, tile=tile)


def test_routes_without_the_fused_kernel_project_the_full_simulation():
    """FSE (its signal is not the recorded sample times a constant) and a
    rank above 16 run the full simulation and project it; the answers equal
    doing that by hand."""
    from torchsim_amd import Subspace
    from torchsim_amd.simulators import FSESimulator, MRFSimulator

    t1, t2 = _cases.random_tissue(300, seed=4)
    values = {"T1": torch.tensor(t1).cuda(), "T2": torch.tensor(t2).cuda()}
    fse = FSESimulator(flip=torch.full((32,), 150.0), ESP=5.0, TR=3000.0, states=10)
    subspace = Subspace.fit(fse.simulate(**values).cpu(), 4)
    coef, jcoef = fse.project(subspace, ("T1", "T2"), **values)
    want, jwant = _unfused(fse, subspace, ("T1", "T2"), values)
    torch.testing.assert_close(coef, want)
    torch.testing.assert_close(jcoef, jwant)

    torch.testing.assert_close(fse.project(subspace, **values), subspace.project(fse.simulate(**values)))

    mrf = MRFSimulator(flip=torch.tensor(_cases.mrf_flips()[:120], dtype=torch.float32), TR=10.0, states=10)
    wide = Subspace.fit(mrf.simulate(**values).cpu(), 20)
    torch.testing.assert_close(mrf.project(wide, **values), wide.project(mrf.simulate(**values)))
    with pytest.raises(ValueError, match="contrasts"):
        mrf.project(subspace, **values)


def test_no_atoms_and_shapes():
    from torchsim_amd import Subspace
    from torchsim_amd.simulators import MRFSimulator

    mrf = MRFSimulator(flip=torch.tensor(_cases.mrf_flips()[:64], dtype=torch.float32), TR=10.0, states=10)
    grid = torch.linspace(300.0, 2000.0, 12).cuda()
    subspace = Subspace.fit(mrf.simulate(T1=grid, T2=grid / 10).cpu(), 3)
    maps = mrf.project(subspace, "T1", T1=grid.reshape(3, 4), T2=grid.reshape(3, 4) / 10)
    assert maps[0].shape == (3, 4, 3) and maps[1].shape == (3, 4, 3)
    empty = mrf.project(subspace, T1=grid[:0], T2=grid[:0])
    assert empty.shape == (0, 3)


def test_subspace_normal_equations_match_the_oracle():
    """Cost, J^T r and J^T J of coefficients: the kernel-projected model and
    Jacobian against the FP64 oracle's, projected in FP64."""
    import oracle
    from tests.test_gpu_fit import _mrf
    from torchsim_amd import Subspace
    from torchsim_amd.recon import FusedModel

    seq = _mrf()
    t1, t2 = _cases.random_tissue(200, seed=23)
    m0 = np.linspace(0.5, 1.0, 200)
    events = oracle.events_of(seq.packed(**seq.bound))
    sig, tan = oracle.real(events, t1=t1, t2=t2, m0=m0, states=10, outputs=880,
                           tangents=[{"t1": 1.0}, {"t2": 1.0}, {"m0": 1.0}])
    model, jac = -sig, -tan  # MRF: 1j * (i m0 a0)
    basis = _basis(model.astype(np.complex128), 8).astype(np.complex64)
    b64 = basis.astype(np.complex128)
    pm, pj = model @ b64, jac @ b64
    rng = np.random.default_rng(4)
    data = (pm + 0.01 * (rng.normal(size=pm.shape) + 1j * rng.normal(size=pm.shape))).astype(np.complex64)
    r = pm - data.astype(np.complex128)
    cost = (np.abs(r) ** 2).sum(-1)
    g = np.einsum("vps,vs->vp", pj.conj(), r).real
    h = np.einsum("vps,vqs->vpq", pj.conj(), pj).real
    op = FusedModel(seq, ("T1", "T2", "M0"), subspace=Subspace(basis=torch.from_numpy(basis),
                                                               singular_values=torch.ones(8)))
    x = torch.tensor(np.stack([t1, t2, m0], -1), dtype=torch.float64, device="cuda")
    normal = op.normal(x, torch.tensor(data).cuda(), None).cpu().numpy()
    assert _cases.rel_err(normal[:, 0], cost) < 1e-3  # cost of a small residual: cancellation-limited
    for p in range(3):
        assert _cases.rel_err(normal[:, 1 + p], g[:, p]) < 1e-3
    index = {(0, 0): 5, (0, 1): 6, (0, 2): 7, (1, 1): 9, (1, 2): 10, (2, 2): 12}
    for (p, q), k in index.items():
        assert _cases.rel_err(normal[:, k], h[:, p, q]) < 1e-4, (p, q)


def test_fits_in_a_fitted_subspace_recover_noiseless_maps():
    """NonlinearLeastSquares with a rank: a basis fitted to simulated
    signals, the measured volume projected onto it, Levenberg-Marquardt on
    the coefficients; from_coefficients on the same coefficients agrees."""
    from tests.test_gpu_fit import _mrf
    from torchsim_amd.estimators import NonlinearLeastSquares
    from torchsim_amd.recon import GaussNewton, TrustRegion

    seq = _mrf()
    t1, t2 = _cases.random_tissue(300, seed=5)
    t1 = np.clip(t1, 300, 2500).astype(np.float32)
    t2 = np.clip(t2, 30, 250).astype(np.float32)
    m0 = np.linspace(0.5, 1.0, 300).astype(np.float32)
    data = seq.simulate(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda(), M0=torch.tensor(m0).cuda())
    fit = NonlinearLeastSquares(seq, bounds={"T1": (100.0, 5000.0), "T2": (10.0, 2000.0), "M0": (0.01, 2.0)},
                                loop=GaussNewton(TrustRegion(), max_iterations=40))
    fit.fit(rank=8, samples=4000, T1=(300.0, 2500.0), T2=(30.0, 250.0), M0=(0.5, 1.0))
    assert fit.subspace.rank == 8 and fit.subspace.retained > 0.999
    maps = fit.map(data)
    for name, truth in (("T1", t1), ("T2", t2), ("M0", m0)):
        err = np.abs(maps[name].cpu().numpy() - truth) / truth
        assert np.median(err) < 1e-3, name
        assert np.percentile(err, 99) < 2e-2, name
    again = fit.from_coefficients(fit.subspace.project(data))
    for name in maps:
        torch.testing.assert_close(again[name], maps[name])
    with pytest.raises(ValueError, match="rank 8"):
        fit.from_coefficients(data)


def test_a_tall_dictionary_fits_on_the_device():
    """Subspace.fit of 20,000 MRF atoms (x 300 contrasts) held on the GPU: the
    Gram route runs there (FP64 GEMM + eigh) and agrees with the FP64 SVD
    oracle of the same atoms."""
    from torchsim_amd import Subspace
    from torchsim_amd.simulators import MRFSimulator

    t1, t2 = _cases.random_tissue(20000, seed=12)
    sim = MRFSimulator(flip=torch.tensor(_cases.mrf_flips()[:300], dtype=torch.float32), TR=10.0, states=10)
    atoms = sim.simulate(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda())
    subspace = Subspace.fit(atoms, 6)
    assert subspace.basis.is_cuda and subspace.basis.dtype == torch.complex64
    basis, values = osub.fit(atoms.cpu().numpy().astype(np.complex128), 6)
    np.testing.assert_allclose(subspace.singular_values.cpu().numpy()[:6], values[:6], rtol=1e-5)
    got = subspace.basis.cpu().numpy().astype(np.complex128)
    np.testing.assert_allclose(got @ got.conj().T, basis @ basis.conj().T, atol=1e-4)
