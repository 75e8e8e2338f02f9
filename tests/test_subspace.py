"""The temporal basis: the FP64 oracle pinned by the reference's own
properties (tests/estimators/test_subspace.py), and the product's Subspace
(torch, any device) against it. CPU only."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import subspace as oracle
from torchsim_amd import Subspace

RANKS = (1, 2, 4, 8)


def _echo_train(count: int = 240, echoes: int = 64, phase: float = 0.7, seed: int = 0) -> np.ndarray:
    """Multi-echo decays exp(-TE/T2) (5 ms spacing, T2 20-270 ms) with a
    partial-recovery weight per atom, all sharing one non-real phase."""
    rng = np.random.default_rng(seed)
    t2 = 20.0 + 250.0 * rng.random(count)
    t1 = 200.0 + 2800.0 * rng.random(count)
    te = 5.0 * np.arange(1, echoes + 1)
    weight = 1.0 - np.exp(-3000.0 / t1)
    return (weight[:, None] * np.exp(-te[None, :] / t2[:, None]) * np.exp(1j * phase)).astype(np.complex128)


def _error(signals, basis, project=oracle.project):
    through = oracle.expand(project(signals, basis), basis)
    return float(np.sum(np.abs(signals - through) ** 2) / np.sum(np.abs(signals) ** 2))


# %% the oracle

@pytest.mark.parametrize("rank", RANKS)
def test_oracle_retained_is_the_round_trip_error(rank):
    """Reference test_subspace.py:29-44."""
    signals = _echo_train()
    basis, values = oracle.fit(signals, rank)
    assert oracle.retained(values, rank) == pytest.approx(1.0 - _error(signals, basis), abs=1e-12)


def test_oracle_full_rank_round_trips_and_is_orthonormal():
    """Reference test_subspace.py:47-64."""
    signals = _echo_train(count=32, echoes=16)
    basis, values = oracle.fit(signals, 16)
    assert oracle.retained(values, 16) == pytest.approx(1.0, abs=1e-12)
    np.testing.assert_allclose(oracle.expand(oracle.project(signals, basis), basis), signals, atol=1e-12)
    np.testing.assert_allclose(basis.conj().T @ basis, np.eye(16), atol=1e-12)


def test_oracle_an_echo_train_needs_few_directions():
    """Reference test_subspace.py:67-73."""
    basis, values = oracle.fit(_echo_train(), 8)
    assert oracle.retained(values, 8) > 0.9999 and basis.shape == (64, 8)


def test_the_round_trip_ignores_each_directions_phase():
    """A direction's phase is arbitrary (a library's SVD picks one), and the
    plain-transpose projection with its conjugate expansion cancels it. A
    planted mistake -- projecting with conj(basis) -- does not, and neither
    does a retained fraction of unsquared singular values."""
    signals = _echo_train()
    basis, values = oracle.fit(signals, 4)
    turned = basis * np.exp(1j * 0.9 * np.arange(1, 5))[None, :]
    assert _error(signals, turned) == pytest.approx(1.0 - oracle.retained(values, 4), abs=1e-12)
    wrong = _error(signals, turned, project=lambda s, b: s @ b.conj())
    assert abs((1.0 - oracle.retained(values, 4)) - wrong) > 1e-3
    unsquared = float(values[:4].sum() / values.sum())
    assert abs((1.0 - unsquared) - _error(signals, basis)) > 1e-4


def test_oracle_ranks_are_checked():
    with pytest.raises(ValueError, match="positive"):
        oracle.fit(_echo_train(), 0)
    with pytest.raises(ValueError, match="exceeds"):
        oracle.fit(_echo_train(count=4, echoes=8), 6)


# %% the product's Subspace against the oracle

@pytest.mark.parametrize("rank", RANKS)
def test_subspace_matches_the_oracle(rank):
    signals = _echo_train()
    subspace = Subspace.fit(torch.from_numpy(signals.astype(np.complex64)), rank)
    basis, values = oracle.fit(signals, rank)
    assert subspace.rank == rank and subspace.contrasts == 64
    assert subspace.basis.dtype == torch.complex64
    np.testing.assert_allclose(subspace.singular_values.numpy(), values, rtol=1e-4, atol=1e-6 * values[0])
    assert subspace.retained == pytest.approx(oracle.retained(values, rank), abs=1e-6)
    # the same projector, whatever phase each direction was given, where the
    # last kept direction stands well above float32 rounding of the signals
    if values[rank - 1] > 1e-3 * values[0]:
        projector = subspace.basis.numpy().astype(np.complex128) @ subspace.basis.numpy().conj().T
        np.testing.assert_allclose(projector, basis @ basis.conj().T, atol=1e-4)
    through = subspace.expand(subspace.project(torch.from_numpy(signals.astype(np.complex64)))).numpy()
    np.testing.assert_allclose(through, oracle.expand(oracle.project(signals, basis), basis), atol=2e-5)


def test_subspace_modes_flattening_and_dtypes():
    signals = torch.from_numpy(_echo_train(count=60, echoes=32).astype(np.complex64))
    flat = Subspace.fit(signals, 5)
    shaped = Subspace.fit(signals.reshape(6, 10, 32), 5)
    torch.testing.assert_close(flat.basis, shaped.basis, atol=1e-6, rtol=0)
    assert flat.modes.shape == (5, 32) and torch.equal(flat.modes, flat.basis.mT)
    # a real basis promotes a complex signal, never narrows it
    real = Subspace.fit(signals.real.contiguous(), 3)
    assert real.basis.dtype == torch.float32
    assert real.project(signals).dtype == torch.complex64
    with pytest.raises(ValueError, match="rank must be positive"):
        Subspace.fit(signals, 0)
    with pytest.raises(ValueError, match="exceeds"):
        Subspace.fit(signals[:4, :8], 6)


def test_a_tall_dictionary_takes_the_gram_route_and_agrees():
    """2,000 atoms over 64 contrasts: the basis from the contrasts' Gram
    matrix (FP64 eigh) against the FP64 SVD oracle: the same singular values
    and the same projector onto the leading directions."""
    signals = _echo_train(count=2000)
    subspace = Subspace.fit(torch.from_numpy(signals.astype(np.complex64)), 6)
    basis, values = oracle.fit(signals, 6)
    assert subspace.singular_values.shape == (64,)
    np.testing.assert_allclose(subspace.singular_values.numpy()[:6], values[:6], rtol=1e-5)
    assert subspace.retained == pytest.approx(oracle.retained(values, 6), abs=1e-6)
    projector = subspace.basis.numpy().astype(np.complex128) @ subspace.basis.numpy().conj().T
    np.testing.assert_allclose(projector, basis @ basis.conj().T, atol=1e-4)
    np.testing.assert_allclose(subspace.basis.numpy().conj().T @ subspace.basis.numpy(), np.eye(6), atol=1e-5)


def test_the_gram_matrix_summed_in_ragged_blocks_is_the_same(monkeypatch):
    """The FP64 Gram matrix is summed one block of atoms at a time; blocks of
    333 atoms over 2,000 (the last one ragged) give the basis of one block."""
    from torchsim_amd import _subspace

    signals = torch.from_numpy(_echo_train(count=2000).astype(np.complex64))
    whole = Subspace.fit(signals, 6)
    monkeypatch.setattr(_subspace, "_GRAM_BLOCK", 333)
    blocked = Subspace.fit(signals, 6)
    top = float(whole.singular_values[0])
    torch.testing.assert_close(blocked.singular_values, whole.singular_values, rtol=0, atol=1e-6 * top)
    projector = lambda b: b.to(torch.complex128) @ b.to(torch.complex128).mH  # noqa: E731
    torch.testing.assert_close(projector(blocked.basis), projector(whole.basis), atol=1e-6, rtol=0)
