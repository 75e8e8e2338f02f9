"""Dictionary matching on the MFMA kernel against the FP64 oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import match as oracle
from tests import _cases

pytestmark = pytest.mark.gpu

# The candidate pass keeps top_k + 1 atoms per voxel from an fp16-split MFMA
# product (~1e-6 of a unit score) and ranks them by an exact FP64 recompute,
# so a ranking agrees with the oracle wherever consecutive oracle scores are
# further apart than this.
TIE = 1e-5


def _matcher(dictionary, parameters=None, top_k=1):
    from torchsim_amd.estimators import DictionaryMatcher

    return DictionaryMatcher(dictionary=dictionary, parameters=parameters, top_k=top_k)


def _check(dictionary, signals, top_k, result):
    d = dictionary.cpu().numpy() if isinstance(dictionary, torch.Tensor) else dictionary
    y = signals.cpu().numpy() if isinstance(signals, torch.Tensor) else signals
    # everything below in FP64 (the inputs as given, widened)
    d = d.astype(np.complex128) if np.iscomplexobj(d) else d.astype(np.float64)
    y = y.astype(np.complex128) if np.iscomplexobj(y) else y.astype(np.float64)
    indices, scores, scales, densities = oracle.match(d, y, top_k)
    decided = oracle.gaps(d, y, top_k) > TIE
    got = result.indices.reshape(-1, top_k).cpu().numpy()
    np.testing.assert_array_equal(got[decided], indices[decided])
    assert decided.mean() > 0.9
    # on every voxel, what was returned is scored exactly
    exact = np.abs(np.einsum("vkn,vn->vk", np.conj(d[got]), y.reshape(-1, d.shape[1])))
    dn = np.maximum(np.linalg.norm(d[got], axis=-1), oracle.EPS32)
    yn = np.maximum(np.linalg.norm(y.reshape(-1, d.shape[1]), axis=-1), oracle.EPS32)
    np.testing.assert_allclose(result.scores.reshape(-1, top_k).cpu().numpy(), exact / (dn * yn[:, None]),
                               rtol=0, atol=1e-6)
    # and never worse than the oracle's best by more than the tie width
    assert np.all(result.scores.reshape(-1, top_k).cpu().numpy()[:, 0] >= scores[:, 0] - TIE)
    np.testing.assert_allclose(result.scales.reshape(-1, top_k).cpu().numpy()[decided], scales[decided],
                               rtol=1e-5, atol=1e-6 * np.abs(scales).max())
    np.testing.assert_allclose(result.densities.reshape(-1, top_k).cpu().numpy()[decided], densities[decided],
                               rtol=1e-5, atol=0)


def test_complex_scaled_atoms_are_recovered():
    """The reference's own known answer (tests/estimators/test_dictionary.py:24-45)."""
    parameter = torch.linspace(0.1, 2.0, 67, device="cuda")[:, None]
    dictionary = torch.cat((torch.exp(1j * parameter), torch.exp(2j * parameter)), dim=-1)
    scale = torch.tensor([2.0 - 0.5j, -0.2 + 1.1j], device="cuda")
    selected = torch.tensor([3, 52], device="cuda")
    signals = scale[:, None] * dictionary[selected]
    result = _matcher(dictionary, parameter, top_k=2).match(signals)
    torch.testing.assert_close(result.indices[:, 0], selected)
    torch.testing.assert_close(result.parameters[:, 0, 0], parameter[selected, 0])
    torch.testing.assert_close(result.scales[:, 0], scale.to(torch.complex64))
    torch.testing.assert_close(result.scores[:, 0], torch.ones(2, device="cuda"))


def test_identity_dictionary_returns_indices_without_parameters():
    """Reference tests/estimators/test_dictionary.py:48-52."""
    dictionary = torch.eye(4, device="cuda")
    found = _matcher(dictionary)(dictionary[[2, 0]])
    torch.testing.assert_close(found, torch.tensor([2, 0], device="cuda"))


@pytest.mark.parametrize("kind", ["real", "complex-real-plane", "complex-imag-plane", "complex"])
@pytest.mark.parametrize("top_k", [1, 2, 3])
def test_random_dictionaries_match_the_oracle(kind, top_k):
    """Every operand mode, ragged sizes (atoms, voxels and contrasts not
    multiples of the tiles), complex and real signals."""
    print(f"case kind={kind} top_k={top_k}", flush=True)
    kinds = ["real", "complex-real-plane", "complex-imag-plane", "complex"]
    rng = np.random.default_rng(10 * kinds.index(kind) + top_k)
    atoms, contrasts, voxels = 3001, 37, 517
    base = rng.standard_normal((atoms, contrasts)).astype(np.float32)
    if kind == "real":
        dictionary = base
    elif kind == "complex-real-plane":
        dictionary = base.astype(np.complex64)
    elif kind == "complex-imag-plane":
        dictionary = (1j * base).astype(np.complex64)
    else:
        dictionary = (base + 1j * rng.standard_normal((atoms, contrasts))).astype(np.complex64)
    pick = rng.integers(0, atoms, voxels)
    gain = (rng.uniform(0.2, 3.0, voxels) * np.exp(2j * np.pi * rng.uniform(size=voxels))).astype(np.complex64)
    noise = 0.3 * (rng.standard_normal((voxels, contrasts)) + 1j * rng.standard_normal((voxels, contrasts)))
    signals = (gain[:, None] * dictionary[pick] + noise).astype(np.complex64)
    result = _matcher(torch.from_numpy(dictionary).cuda(), top_k=top_k).match(torch.from_numpy(signals).cuda())
    _check(dictionary, signals, top_k, result)


def test_real_signals_against_a_real_dictionary_give_real_scales():
    rng = np.random.default_rng(5)
    dictionary = rng.standard_normal((700, 20)).astype(np.float32)
    signals = (dictionary[rng.integers(0, 700, 33)] * 1.7 + 0.1 * rng.standard_normal((33, 20))).astype(np.float32)
    result = _matcher(torch.from_numpy(dictionary).cuda(), top_k=2).match(torch.from_numpy(signals).cuda())
    assert not torch.is_complex(result.scales)
    _check(dictionary, signals, 2, result)


def test_equal_atoms_rank_the_lower_index_first():
    rng = np.random.default_rng(2)
    base = (rng.standard_normal((300, 16)) + 1j * rng.standard_normal((300, 16))).astype(np.complex64)
    twin = np.concatenate([base, base[:40]])  # atoms 300..339 repeat atoms 0..39
    result = _matcher(torch.from_numpy(twin).cuda(), top_k=2).match(torch.from_numpy(base[[1, 3, 39]]).cuda())
    np.testing.assert_array_equal(result.indices.cpu().numpy(), [[1, 301], [3, 303], [39, 339]])


def test_mrf_dictionary_recovers_the_tissue_of_noisy_voxels():
    """An MRF dictionary from the simulator (its signal, -M0 a0, is real: the
    one-plane operand), voxels drawn from its atoms with a complex receive
    gain and noise."""
    from torchsim_amd.simulators import MRFSimulator

    t1 = torch.tensor(np.geomspace(100, 5000, 60), dtype=torch.float32)
    t2 = torch.tensor(np.geomspace(10, 2000, 50), dtype=torch.float32)
    grid1, grid2 = torch.meshgrid(t1, t2, indexing="ij")
    keep = grid2 < grid1
    t1s, t2s = grid1[keep].cuda(), grid2[keep].cuda()
    sim = MRFSimulator(flip=torch.tensor(_cases.benchmark_flips(880), dtype=torch.float32), TR=10.0, states=20)
    atoms = sim.simulate(T1=t1s, T2=t2s)
    parameters = torch.stack([t1s, t2s], dim=-1)
    rng = np.random.default_rng(3)
    pick = torch.from_numpy(rng.integers(0, atoms.shape[0], 2000)).cuda()
    gain = torch.from_numpy((rng.uniform(0.5, 1.0, 2000) * np.exp(2j * np.pi * rng.uniform(size=2000)))
                            .astype(np.complex64)).cuda()
    sigma = float(atoms.abs().max()) / 50.0
    noise = sigma * torch.randn(2000, atoms.shape[1], dtype=torch.complex64, device="cuda",
                                generator=torch.Generator(device="cuda").manual_seed(3))
    signals = gain[:, None] * atoms[pick] + noise
    matcher = _matcher(atoms, parameters)
    assert matcher._mode == 1
    result = matcher.match(signals)
    _check(atoms, signals, 1, result)
    assert (result.indices[:, 0] == pick).float().mean() > 0.3


def test_no_voxels_and_shapes():
    dictionary = torch.randn(10, 8, dtype=torch.complex64, device="cuda")
    matcher = _matcher(dictionary, torch.arange(10.0), top_k=3)
    empty = matcher.match(torch.zeros(0, 8, dtype=torch.complex64, device="cuda"))
    assert empty.indices.shape == (0, 3) and empty.parameters.shape == (0, 3, 1)
    volume = matcher.match(dictionary[torch.tensor([[1, 2], [3, 4]], device="cuda")])
    assert volume.indices.shape == (2, 2, 3) and volume.parameters.shape == (2, 2, 3, 1)
    torch.testing.assert_close(volume.indices[..., 0], torch.tensor([[1, 2], [3, 4]], device="cuda"))
    with pytest.raises(ValueError):
        matcher.match(torch.zeros(3, 7, device="cuda"))
    with pytest.raises(ValueError):
        _matcher(dictionary, top_k=11)  # more candidates than atoms
    assert _matcher(dictionary, top_k=4).match(dictionary[:2]).indices.shape == (2, 4)


def test_matching_in_a_fitted_subspace():
    """fit(rank=): the atoms held as coefficients of a basis fitted to them,
    voxels projected onto it by match(); the ranking is the oracle's on those
    coefficients, and from_coefficients on the projected voxels agrees."""
    from torchsim_amd.simulators import MRFSimulator

    t1 = torch.tensor(np.geomspace(100, 5000, 40), dtype=torch.float32)
    t2 = torch.tensor(np.geomspace(10, 2000, 30), dtype=torch.float32)
    grid1, grid2 = torch.meshgrid(t1, t2, indexing="ij")
    keep = grid2 < grid1
    t1s, t2s = grid1[keep].cuda(), grid2[keep].cuda()
    sim = MRFSimulator(flip=torch.tensor(_cases.benchmark_flips(880), dtype=torch.float32), TR=10.0, states=10)
    atoms = sim.simulate(T1=t1s, T2=t2s)
    rng = np.random.default_rng(8)
    pick = torch.from_numpy(rng.integers(0, atoms.shape[0], 700)).cuda()
    gain = torch.from_numpy((rng.uniform(0.5, 1.0, 700) * np.exp(2j * np.pi * rng.uniform(size=700)))
                            .astype(np.complex64)).cuda()
    noise = (float(atoms.abs().max()) / 50.0) * torch.randn(700, 880, dtype=torch.complex64, device="cuda",
                                                             generator=torch.Generator(device="cuda").manual_seed(8))
    signals = gain[:, None] * atoms[pick] + noise
    from torchsim_amd.estimators import DictionaryMatcher

    matcher = DictionaryMatcher(top_k=2).fit(atoms.cpu(), torch.stack([t1s, t2s], -1), rank=10)
    assert matcher.subspace.rank == 10 and matcher.dictionary.shape == (atoms.shape[0], 10)
    result = matcher.match(signals)
    coefficients = matcher.subspace.project(signals)
    _check(matcher.dictionary, coefficients, 2, result)
    again = matcher.from_coefficients(coefficients)
    torch.testing.assert_close(again.indices, result.indices, rtol=0, atol=0)
    torch.testing.assert_close(again.scores, result.scores, rtol=0, atol=0)
    assert (result.indices[:, 0] == pick).float().mean() > 0.3
    with pytest.raises(ValueError, match="rank 10"):
        matcher.from_coefficients(signals)


@pytest.mark.parametrize("kind", ["real", "complex"])
@pytest.mark.parametrize("top_k", [4, 9])
def test_top_k_beyond_the_fused_kernel(kind, top_k):
    """top_k above 3 runs on a library GEMM per chunk with an FP64 refine
    (reference _dictionary.py:120-136, 289-290): same checks as the fused
    path, chunks smaller than the dictionary and the voxel set."""
    from torchsim_amd.estimators import DictionaryMatcher

    rng = np.random.default_rng(70 + top_k)
    atoms, contrasts, voxels = 2003, 33, 301
    dictionary = rng.standard_normal((atoms, contrasts)).astype(np.float32)
    if kind == "complex":
        dictionary = (dictionary + 1j * rng.standard_normal((atoms, contrasts))).astype(np.complex64)
    pick = rng.integers(0, atoms, voxels)
    gain = (rng.uniform(0.2, 3.0, voxels) * np.exp(2j * np.pi * rng.uniform(size=voxels))).astype(np.complex64)
    noise = 0.3 * (rng.standard_normal((voxels, contrasts)) + 1j * rng.standard_normal((voxels, contrasts)))
    signals = (gain[:, None] * dictionary[pick] + noise).astype(np.complex64)
    matcher = DictionaryMatcher(dictionary=torch.from_numpy(dictionary).cuda(), top_k=top_k,
                                query_chunk_size=128, dictionary_chunk_size=700)
    result = matcher.match(torch.from_numpy(signals).cuda())
    _check(dictionary, signals, top_k, result)


def _fingerprints():
    """A fingerprinting dictionary over a T1/T2 grid (as the reference's
    grouped-matching tests build it) and noisy voxels off it."""
    from torchsim_amd.simulators import MRFSimulator

    generator = torch.Generator().manual_seed(0)
    flip = 10.0 + 50.0 * torch.rand(200, generator=generator)
    grid = torch.cartesian_prod(torch.linspace(200.0, 3000.0, 60), torch.linspace(20.0, 300.0, 30))
    atoms = MRFSimulator(TR=10.0, TI=20.0).simulate(flip=flip, T1=grid[:, 0].cuda(), T2=grid[:, 1].cuda())
    generator = torch.Generator().manual_seed(7)
    picked = atoms[::37].cpu()
    noise = torch.randn(picked.shape, generator=generator, dtype=picked.dtype)
    return atoms, grid.cuda(), (picked + 0.01 * float(atoms.abs().max()) * noise).cuda()


@pytest.mark.parametrize("top_k", [1, 3])
def test_grouped_matching_matches_the_oracle(top_k):
    """groups=24 at the default prune: the grouping is the oracle's (same
    members), and each voxel's answer is the oracle's grouped search."""
    from torchsim_amd.estimators import DictionaryMatcher

    atoms, grid, measured = _fingerprints()
    matcher = DictionaryMatcher(dictionary=atoms, parameters=grid, groups=24, top_k=top_k)
    members, _ = oracle.grouping(atoms.cpu().numpy(), 24)
    got = [set(m.cpu().tolist()) for m in matcher.grouping.members]
    assert got == [set(m.tolist()) for m in members]
    result = matcher.match(measured)
    indices, scores = oracle.match_grouped(atoms.cpu().numpy(), measured.cpu().numpy(), 24, matcher.prune, top_k)
    np.testing.assert_array_equal(result.indices.cpu().numpy(), indices)
    np.testing.assert_allclose(result.scores.cpu().numpy(), scores, rtol=0, atol=1e-6)


def test_grouped_matching_finds_what_direct_matching_finds():
    """Reference tests/estimators/test_grouped.py:81-100: most voxels get the
    same parameters, and where they differ the error is a grid step."""
    from torchsim_amd.estimators import DictionaryMatcher

    atoms, grid, measured = _fingerprints()
    direct = DictionaryMatcher(dictionary=atoms, parameters=grid)(measured)
    grouped = DictionaryMatcher(dictionary=atoms, parameters=grid, groups=24)(measured)
    assert float((direct == grouped).all(-1).float().mean()) > 0.8
    for column in range(grid.shape[-1]):
        assert float(((grouped[:, column] - direct[:, column]) / direct[:, column]).abs().mean()) < 0.01
    wide = DictionaryMatcher(dictionary=atoms, parameters=grid, groups=24, prune=0.999999)(measured)
    torch.testing.assert_close(wide, direct)  # keeping every group is the exhaustive search


def test_grouped_matching_in_a_fitted_subspace():
    """Grouping happens inside the global basis (reference test_grouped.py:
    215-260): the atoms are clustered as rank-8 coefficients."""
    from torchsim_amd.estimators import DictionaryMatcher

    atoms, grid, measured = _fingerprints()
    matcher = DictionaryMatcher(groups=24).fit(atoms, grid, rank=8)
    assert matcher.grouping.representative.shape == (24, 8)
    direct = DictionaryMatcher().fit(atoms, grid, rank=8)(measured)
    grouped = matcher(measured)
    assert float((direct == grouped).all(-1).float().mean()) > 0.8
