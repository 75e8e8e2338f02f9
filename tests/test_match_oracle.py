"""Pins of the FP64 dictionary-matching oracle (CPU only)."""

from __future__ import annotations

import numpy as np

from oracle import match as oracle


def scaled_atoms_case():
    """The reference's known answer (tests/estimators/test_dictionary.py:24-45)."""
    parameter = np.linspace(0.1, 2.0, 67)[:, None]
    dictionary = np.concatenate((np.exp(1j * parameter), np.exp(2j * parameter)), axis=-1)
    scale = np.array([2.0 - 0.5j, -0.2 + 1.1j])
    selected = np.array([3, 52])
    signals = scale[:, None] * dictionary[selected]
    return dictionary, signals, scale, selected


def test_scaled_atoms_are_recovered_with_score_one_and_their_scale():
    dictionary, signals, scale, selected = scaled_atoms_case()
    indices, scores, scales, densities = oracle.match(dictionary, signals, top_k=2)
    np.testing.assert_array_equal(indices[:, 0], selected)
    np.testing.assert_allclose(scores[:, 0], 1.0, rtol=0, atol=1e-12)
    np.testing.assert_allclose(scales[:, 0], scale, rtol=1e-12)
    np.testing.assert_allclose(densities[:, 0], np.abs(scale), rtol=1e-12)
    assert np.all(scores[:, 1] < scores[:, 0])


def test_identity_dictionary_returns_the_rows_indices():
    """Reference tests/estimators/test_dictionary.py:48-52."""
    dictionary = np.eye(4)
    indices, *_ = oracle.match(dictionary, dictionary[[2, 0]])
    np.testing.assert_array_equal(indices[:, 0], [2, 0])


def test_score_is_blind_to_the_voxel_phase_and_a_real_part_score_is_not():
    """A complex measurement of a real model (reference test_dictionary.py:55-99):
    the oracle's score is phase invariant; scoring the real part alone, the
    plausible mistake, loses the match at a quarter turn."""
    rng = np.random.default_rng(0)
    t2 = np.linspace(20.0, 300.0, 128)
    te = np.linspace(10.0, 200.0, 16)
    atoms = np.exp(-te[None, :] / t2[:, None])
    truth = np.arange(0, 128, 7)
    noise = 0.002 * (rng.standard_normal((truth.size, 16)) + 1j * rng.standard_normal((truth.size, 16)))
    in_phase, in_phase_scores, _, _ = oracle.match(atoms, atoms[truth] + noise)
    assert np.mean(np.abs(t2[in_phase[:, 0]] - t2[truth])) < np.diff(t2).max()
    for turn in (0.15, 0.25, 0.5):
        phase = np.exp(2j * np.pi * turn)
        found, scores, _, _ = oracle.match(atoms, (atoms[truth] + noise) * phase)
        np.testing.assert_array_equal(found, in_phase)
        np.testing.assert_allclose(scores, in_phase_scores, rtol=1e-12)
    # planted mistake: only the real part of the measurement
    quarter = atoms[truth] * 1j + noise
    wrong, *_ = oracle.match(atoms, quarter.real)
    assert np.mean(wrong[:, 0] != truth) > 0.5


def test_scale_is_the_least_squares_projection():
    rng = np.random.default_rng(1)
    dictionary = rng.standard_normal((50, 24)) + 1j * rng.standard_normal((50, 24))
    signals = rng.standard_normal((9, 24)) + 1j * rng.standard_normal((9, 24))
    indices, scores, scales, densities = oracle.match(dictionary, signals, top_k=3)
    for v in range(9):
        for r in range(3):
            d = dictionary[indices[v, r]]
            residual = signals[v] - scales[v, r] * d
            assert abs(np.vdot(d, residual)) < 1e-10 * np.linalg.norm(d) * np.linalg.norm(signals[v])
            # Cauchy-Schwarz: a score is a cosine
            assert 0.0 <= scores[v, r] <= 1.0 + 1e-12
            np.testing.assert_allclose(densities[v, r], abs(scales[v, r]), rtol=1e-10)
    # best first, ties broken by the lower index
    assert np.all(np.diff(scores, axis=1) <= 0)
    twin = np.concatenate([dictionary, dictionary[:5]])
    again, *_ = oracle.match(twin, dictionary[[1, 3]], top_k=2)
    np.testing.assert_array_equal(again, [[1, 51], [3, 53]])
