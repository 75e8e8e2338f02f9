"""Dictionary matching on the MFMA kernel (reference estimators/_dictionary.py).

The reference scores a chunk of voxels against a chunk of atoms with a torch
matmul, keeps ``torch.topk`` of each chunk and merges the chunks
(``_dictionary.py:333-387``, ``_grouped.py:219-245``). Here one launch scores
every voxel against every atom on the matrix cores and keeps the best
candidates in registers; the survivors are recomputed exactly in FP64 and
ranked by that (``tsim_match``). The score matrix never exists, so the chunk
sizes only bound the caller's expectations and are validated, not used.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any

import torch

from .. import _native
from ._grouped import Grouping, correlate

#: Candidates the fused matcher keeps per voxel beyond top_k; above this
#: top_k the search is a library GEMM per chunk with an FP64 refine.
_FUSED_TOP_K = 3


@dataclass(frozen=True)
class MatchResult:
    """Best atoms per voxel (reference _dictionary.py:17-51).

    ``parameters`` ``(..., k, p)`` (None without parameter values),
    ``indices`` ``(..., k)`` int64, ``scores`` ``(..., k)`` the normalized
    inner products, ``scales`` ``(..., k)`` the complex least-squares scale of
    the measurement on each atom, ``densities`` ``(..., k)`` its size.
    """

    parameters: torch.Tensor | None
    indices: torch.Tensor
    scores: torch.Tensor
    scales: torch.Tensor
    densities: torch.Tensor


class DictionaryMatcher:
    """Match signals to the atoms of a dictionary by normalized complex inner
    products (reference _dictionary.py:54-146).

    ``dictionary`` ``(atoms, contrasts)`` real or complex; ``parameters``
    ``(atoms, p)`` what each atom stands for; ``top_k`` candidates per voxel
    (up to 3 on the fused MFMA kernel, more through a library GEMM per chunk);
    ``groups`` clusters the dictionary and searches each voxel only in the
    groups whose representative scores within ``prune`` of its best
    (reference estimators/_grouped.py, Cauley et al. 2015).
    """

    def __init__(self, acquisition: Any = None, *, dictionary: Any = None, parameters: Any = None,
                 query_chunk_size: int = 4096, dictionary_chunk_size: int = 16384, top_k: int = 1,
                 groups: int | None = None, prune: float = 5e-3) -> None:
        if query_chunk_size < 1 or dictionary_chunk_size < 1:
            raise ValueError("chunk sizes must be positive")
        if top_k < 1:
            raise ValueError("top_k must be at least one")
        if groups is not None and groups < 1:
            raise ValueError(f"groups must be positive, got {groups}")
        if not 0.0 <= prune < 1.0:
            raise ValueError(f"prune must be in [0, 1), got {prune}")
        self.acquisition = acquisition
        self.query_chunk_size = int(query_chunk_size)
        self.dictionary_chunk_size = int(dictionary_chunk_size)
        self.top_k = int(top_k)
        self.groups = None if groups is None else int(groups)
        self.prune = float(prune)
        self._grouping: Grouping | None = None
        self._group_matchers: list = []
        self.dictionary = torch.empty(0)
        self.dictionary_power = torch.empty(0)
        self.parameter_values = torch.empty(0)
        self._operand: tuple[torch.Tensor, torch.Tensor] | None = None
        self._norm: torch.Tensor | None = None
        self._mode = 0
        self._subspace: Any = None
        if dictionary is not None:
            self._adopt(dictionary, parameters)

    @property
    def fitted(self) -> bool:
        return self.dictionary.numel() != 0

    @property
    def grouping(self) -> Grouping | None:
        """The clustered dictionary when ``groups`` was given."""
        return self._grouping

    @property
    def subspace(self) -> Any:
        """The temporal basis the dictionary is held in, if any."""
        return self._subspace

    def fit(self, signals: Any, parameters: Any = None, *, rank: int | None = None,
            subspace: Any = None) -> DictionaryMatcher:
        """Adopt ``signals`` ``(atoms, contrasts)`` as the dictionary. With
        ``rank`` a basis of that rank is fitted to them first, with
        ``subspace`` one is given; either way the atoms are held as their
        coefficients and :meth:`match` projects the voxels onto the same
        basis (reference estimators/_mapping.py:118-255, 508-520)."""
        if rank is not None and subspace is not None:
            raise ValueError("give a rank to fit a basis here, or a subspace to work in one, not both")
        if rank is not None:
            from .._subspace import Subspace

            subspace = Subspace.fit(torch.as_tensor(signals), rank)
        if subspace is not None:
            signals = subspace.project(torch.as_tensor(signals))
        self._adopt(signals, parameters)
        self._subspace = subspace
        return self

    @torch.no_grad()
    def from_coefficients(self, coefficients: Any) -> MatchResult:
        """Match coefficients already in :attr:`subspace` without projecting
        them again (reference estimators/_mapping.py:453-498)."""
        if self._subspace is None:
            raise RuntimeError("this matcher has no subspace, so there are no coefficients to read; "
                               "give fit a rank or a subspace")
        values = torch.as_tensor(coefficients)
        if values.shape[-1] != self._subspace.rank:
            raise ValueError(f"the basis has rank {self._subspace.rank}, got {values.shape[-1]} coefficients")
        return self._match(values)

    def _adopt(self, dictionary: Any, parameters: Any) -> None:
        """Keep these atoms and lay them out for the kernel
        (reference _dictionary.py:284-306)."""
        dictionary = torch.as_tensor(dictionary)
        if dictionary.ndim != 2 or dictionary.shape[0] < 1:
            raise ValueError("dictionary must have shape (atoms, contrasts)")
        if self.top_k > dictionary.shape[0]:
            raise ValueError("top_k must be between one and the atom count")
        if not torch.is_floating_point(dictionary) and not torch.is_complex(dictionary):
            dictionary = dictionary.to(torch.float32)
        if dictionary.is_cuda:
            device = dictionary.device
        elif torch.cuda.is_available():
            device = torch.device("cuda", torch.cuda.current_device())
        else:
            device = torch.device("cpu")
        _native.require_device(device)
        is_complex = torch.is_complex(dictionary)
        dictionary = dictionary.to(device=device, dtype=torch.complex64 if is_complex else torch.float32).contiguous()
        atoms, contrasts = dictionary.shape
        lib = _native.load()
        stream = _native.stream_of(device)
        mode = 0
        with torch.cuda.device(device):
            if is_complex:
                flags = torch.zeros(1, dtype=torch.int32, device=device)
                _native.check(lib.tsim_match_planes(dictionary.data_ptr(), atoms, contrasts, flags.data_ptr(), stream))
                planes = int(flags.item())
                mode = {0: 1, 1: 1, 2: 2}.get(planes, 3)
            atoms_pad, k_pad = ctypes.c_int64(), ctypes.c_int64()
            _native.check(lib.tsim_match_operand_size(atoms, contrasts, mode, ctypes.byref(atoms_pad),
                                                      ctypes.byref(k_pad)))
            hi = torch.empty(atoms_pad.value * k_pad.value, dtype=torch.float16, device=device)
            lo = torch.empty_like(hi)
            norm = torch.empty(atoms, dtype=torch.float64, device=device)
            _native.check(lib.tsim_match_layout_atoms(dictionary.data_ptr(), int(is_complex), atoms, contrasts, mode,
                                                      norm.data_ptr(), hi.data_ptr(), lo.data_ptr(), stream))
        self.dictionary = dictionary
        self._operand = (hi, lo)
        self._norm = norm
        self._mode = mode
        eps = torch.finfo(torch.float32).eps
        self.dictionary_power = norm.clamp_min(eps).square().to(torch.float32)
        self.parameter_values = _prepare_parameters(parameters, dictionary)
        self._grouping = None
        self._group_matchers = []
        if self.groups is not None:
            self._grouping = Grouping.fit(dictionary, min(self.groups, atoms))
            for members in self._grouping.members:
                inner = DictionaryMatcher(top_k=min(self.top_k, int(members.numel())))
                inner._adopt(dictionary[members], None)
                self._group_matchers.append(inner)

    @torch.no_grad()
    def match(self, signals: Any) -> MatchResult:
        """The top atoms, scores, scales, densities and parameters of every
        voxel (reference _dictionary.py:308-331)."""
        if not self.fitted:
            raise RuntimeError("the matcher has no dictionary to match against")
        signals = torch.as_tensor(signals)
        if self._subspace is not None:
            if signals.shape[-1] != self._subspace.contrasts:
                raise ValueError("signal and dictionary contrast counts differ")
            signals = self._subspace.project(signals)
        return self._match(signals)

    def _match(self, signals: torch.Tensor) -> MatchResult:
        if signals.shape[-1] != self.dictionary.shape[-1]:
            raise ValueError("signal and dictionary contrast counts differ")
        sample_shape = signals.shape[:-1]
        contrasts = self.dictionary.shape[-1]
        device = self.dictionary.device
        signals_complex = torch.is_complex(signals)
        flat = signals.reshape(-1, contrasts).to(
            device=device, dtype=torch.complex64 if signals_complex else torch.float32).contiguous()
        if self._grouping is not None:
            indices, scores, scales, densities = self._match_groups(flat)
        else:
            indices, scores, scales, densities = self._match_flat(flat, self.top_k)
        if not signals_complex and not torch.is_complex(self.dictionary):
            scales = scales.real.contiguous()
        k = indices.shape[-1]
        shape = (*sample_shape, k)
        parameters = None
        if self.parameter_values.numel():
            parameters = self.parameter_values[indices].reshape(*shape, self.parameter_values.shape[-1])
        return MatchResult(parameters=parameters, indices=indices.reshape(shape), scores=scores.reshape(shape),
                           scales=scales.reshape(shape), densities=densities.reshape(shape))

    def _match_flat(self, flat: torch.Tensor, k: int) -> tuple[torch.Tensor, ...]:
        """(indices, scores, complex scales, densities), each (voxels, k), of
        ``flat`` (voxels, contrasts) on the dictionary's device."""
        if k > _FUSED_TOP_K:
            return self._match_gemm(flat, k)
        device = flat.device
        voxels, contrasts = flat.shape
        indices = torch.empty((voxels, k), dtype=torch.int64, device=device)
        scores = torch.empty((voxels, k), dtype=torch.float32, device=device)
        scales = torch.empty((voxels, k), dtype=torch.complex64, device=device)
        densities = torch.empty((voxels, k), dtype=torch.float32, device=device)
        hi, lo = self._operand
        lib = _native.load()
        with torch.cuda.device(device):
            _native.check(lib.tsim_match(
                hi.data_ptr(), lo.data_ptr(), self._mode, self.dictionary.shape[0], self.dictionary.data_ptr(),
                int(torch.is_complex(self.dictionary)), self._norm.data_ptr(), flat.data_ptr(),
                int(torch.is_complex(flat)), voxels, contrasts, k, 0, indices.data_ptr(), scores.data_ptr(),
                scales.data_ptr(), densities.data_ptr(), _native.stream_of(device)))
        return indices, scores, scales, densities

    def _match_gemm(self, flat: torch.Tensor, k: int) -> tuple[torch.Tensor, ...]:
        """top_k above what the fused kernel keeps: normalized scores by a
        library GEMM per (voxel chunk, atom chunk), the best k + 1 of each
        kept (FP32 ranking), then those candidates scored exactly in FP64 and
        the best k by exact score (ties: the lower atom index) with their
        least-squares scales and densities (reference _dictionary.py:333-431)."""
        eps = torch.finfo(torch.float32).eps
        atoms = self.dictionary.shape[0]
        keep = min(k + 1, atoms)
        unit_atoms = self.dictionary / self._norm.clamp_min(eps).to(torch.float32)[:, None]
        picks = []
        for query in flat.split(self.query_chunk_size):
            unit = query / torch.linalg.vector_norm(query, dim=-1, keepdim=True).clamp_min(eps)
            best_s = torch.empty((query.shape[0], 0), dtype=torch.float32, device=flat.device)
            best_i = torch.empty((query.shape[0], 0), dtype=torch.int64, device=flat.device)
            for start in range(0, atoms, self.dictionary_chunk_size):
                block = correlate(unit, unit_atoms[start:start + self.dictionary_chunk_size])
                s_chunk, i_chunk = torch.topk(block, min(keep, block.shape[-1]), dim=-1)
                best_s = torch.cat((best_s, s_chunk), -1)
                best_i = torch.cat((best_i, i_chunk + start), -1)
                best_s, order = torch.topk(best_s, min(keep, best_s.shape[-1]), dim=-1)
                best_i = torch.gather(best_i, -1, order)
            picks.append(best_i)
        candidates = torch.cat(picks)
        wide = torch.complex128 if torch.is_complex(self.dictionary) or torch.is_complex(flat) else torch.float64
        y = flat.to(wide)
        d = self.dictionary.to(wide)[candidates]                       # (voxels, keep, contrasts)
        inner = torch.einsum("vkc,vc->vk", d.conj(), y)                 # <d, y>
        dn = self._norm[candidates].clamp_min(eps)
        yn = torch.linalg.vector_norm(y, dim=-1).clamp_min(eps)
        exact = inner.abs() / (yn[:, None] * dn)
        # best k by exact score, the lower index first among equals
        order = torch.argsort(candidates, dim=-1, stable=True)
        exact_sorted = torch.gather(exact, -1, order)
        rank = torch.argsort(-exact_sorted, dim=-1, stable=True)[:, :k]
        chosen = torch.gather(order, -1, rank)
        indices = torch.gather(candidates, -1, chosen)
        scores = torch.gather(exact, -1, chosen)
        scales = torch.gather(inner, -1, chosen) / torch.gather(dn, -1, chosen) ** 2
        densities = yn[:, None] * scores / torch.gather(dn, -1, chosen)
        return (indices, scores.to(torch.float32), scales.to(torch.complex64), densities.to(torch.float32))

    def _match_groups(self, flat: torch.Tensor) -> tuple[torch.Tensor, ...]:
        """Search each voxel in the groups that survive the prune against the
        representatives (reference _grouped.py:219-245), each group on the
        fused matcher, and keep the best top_k across them (higher exact
        score, then lower atom index)."""
        eps = torch.finfo(torch.float32).eps
        voxels, k = flat.shape[0], self.top_k
        device = flat.device
        unit = flat / torch.linalg.vector_norm(flat, dim=-1, keepdim=True).clamp_min(eps)
        keep = self._grouping.survivors(unit, self.prune)
        indices = torch.zeros((voxels, k), dtype=torch.int64, device=device)
        scores = torch.full((voxels, k), -1.0, dtype=torch.float32, device=device)
        scales = torch.zeros((voxels, k), dtype=torch.complex64, device=device)
        densities = torch.zeros((voxels, k), dtype=torch.float32, device=device)
        for group in keep.any(0).nonzero(as_tuple=True)[0].tolist():
            rows = keep[:, group].nonzero(as_tuple=True)[0]
            inner = self._group_matchers[group]
            g_idx, g_score, g_scale, g_density = inner._match_flat(flat[rows].contiguous(), inner.top_k)
            g_idx = self._grouping.members[group][g_idx]
            cand_i = torch.cat((indices[rows], g_idx), -1)
            cand_s = torch.cat((scores[rows], g_score), -1)
            cand_c = torch.cat((scales[rows], g_scale.to(torch.complex64)), -1)
            cand_d = torch.cat((densities[rows], g_density), -1)
            order = torch.argsort(cand_i, dim=-1, stable=True)
            by_score = torch.argsort(-torch.gather(cand_s, -1, order), dim=-1, stable=True)[:, :k]
            pick = torch.gather(order, -1, by_score)
            indices[rows] = torch.gather(cand_i, -1, pick)
            scores[rows] = torch.gather(cand_s, -1, pick)
            scales[rows] = torch.gather(cand_c, -1, pick)
            densities[rows] = torch.gather(cand_d, -1, pick)
        return indices, scores, scales, densities

    def __call__(self, signals: Any) -> torch.Tensor:
        """Best parameter values, or atom indices without parameters
        (reference _dictionary.py:216-236)."""
        result = self.match(signals)
        if result.parameters is None:
            return result.indices[..., 0]
        return result.parameters[..., 0, :]


def _prepare_parameters(parameters: Any, dictionary: torch.Tensor) -> torch.Tensor:
    """(atoms, p) parameter values on the dictionary's device
    (reference _dictionary.py:437-452)."""
    if parameters is None:
        return torch.empty(0, dtype=torch.float32, device=dictionary.device)
    output = torch.as_tensor(parameters, dtype=dictionary.real.dtype, device=dictionary.device)
    if output.ndim == 1:
        output = output[:, None]
    if output.ndim != 2 or output.shape[0] != dictionary.shape[0]:
        raise ValueError("parameters must have shape (atoms, parameters)")
    return output
