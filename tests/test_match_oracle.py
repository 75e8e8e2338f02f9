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


# %% grouped matching (reference estimators/_grouped.py, tests/estimators/test_grouped.py)

def _decays(count=600, contrasts=40, seed=3):
    """A smooth dictionary: multi-echo decays over T2 with a recovery weight
    over T1, complex with a common phase (nearby atoms nearly parallel)."""
    rng = np.random.default_rng(seed)
    t2 = np.sort(rng.uniform(10, 300, count))
    t1 = rng.uniform(300, 3000, count)
    te = 5.0 * np.arange(1, contrasts + 1)
    weight = 1 - 2 * np.exp(-te[None, :] * 20 / t1[:, None])
    return (weight * np.exp(-te[None, :] / t2[:, None]) * np.exp(0.4j)).astype(np.complex128)


def test_grouping_is_a_partition_of_the_asked_sizes():
    """Reference test_grouped.py:37-58: every atom in exactly one group; all
    groups but the last hold ceil(atoms / count)."""
    d = _decays()
    members, representative = oracle.grouping(d, 24)
    assert len(members) == 24 and representative.shape == (24, d.shape[1])
    np.testing.assert_array_equal(np.sort(np.concatenate(members)), np.arange(d.shape[0]))
    share = -(-d.shape[0] // 24)
    assert {m.size for m in members[:-1]} == {share} and 0 < members[-1].size <= share


def test_grouping_recovers_planted_clusters():
    """Atoms drawn around 6 orthogonal directions (indices interleaved) fall
    into exactly those 6 groups, each seeded by its first unassigned atom,
    and each representative is its cluster's direction."""
    rng = np.random.default_rng(4)
    directions = np.linalg.qr(rng.normal(size=(30, 6)) + 1j * rng.normal(size=(30, 6)))[0].T  # (6, 30)
    label = np.arange(120) % 6
    d = directions[label] * (1 + 0.1 * rng.random(120))[:, None] + 0.01 * rng.normal(size=(120, 30))
    members, representative = oracle.grouping(d, 6)
    for m in members:
        assert np.unique(label[m]).size == 1
    assert [int(label[m[0]]) for m in members] == [0, 1, 2, 3, 4, 5]  # seeds: first unassigned atom
    for g, m in enumerate(members):
        assert abs(np.vdot(directions[label[m[0]]], representative[g])) > 0.999


def test_keeping_every_group_is_the_exhaustive_search_and_pruning_keeps_fewer():
    d = _decays()
    rng = np.random.default_rng(5)
    y = d[::17] * (0.7 * np.exp(1j * rng.uniform(0, 6, (d[::17].shape[0], 1)))) + 1e-3 * rng.normal(size=(36, 40))
    members, representative = oracle.grouping(d, 24)
    every = oracle.survivors(representative, y, 0.999999)
    assert every.all()
    idx_all, sc_all = oracle.match_grouped(d, y, 24, 0.999999, top_k=3)
    idx, sc, _, _ = oracle.match(d, y, top_k=3)
    np.testing.assert_array_equal(idx_all, idx)
    np.testing.assert_allclose(sc_all, sc, rtol=1e-12)
    tight = oracle.survivors(representative, y, 0.0).sum(1)
    loose = oracle.survivors(representative, y, 0.2).sum(1)
    assert tight.min() >= 1 and loose.mean() > tight.mean()
    assert oracle.survivors(representative, y, 5e-3).sum(1).mean() < 0.25 * 24

