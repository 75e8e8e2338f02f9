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

_MAX_TOP_K = 3


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
    (1 to 3 here). ``groups`` (the reference's clustered search) is not built:
    this matcher always searches the whole dictionary, which finds the best
    atom that a grouped search approximates.
    """

    def __init__(self, acquisition: Any = None, *, dictionary: Any = None, parameters: Any = None,
                 query_chunk_size: int = 4096, dictionary_chunk_size: int = 16384, top_k: int = 1,
                 groups: int | None = None, prune: float = 5e-3) -> None:
        if query_chunk_size < 1 or dictionary_chunk_size < 1:
            raise ValueError("chunk sizes must be positive")
        if top_k < 1:
            raise ValueError("top_k must be at least one")
        if top_k > _MAX_TOP_K:
            raise NotImplementedError(f"the fused matcher keeps at most {_MAX_TOP_K} candidates, got top_k={top_k}")
        if groups is not None and groups < 1:
            raise ValueError(f"groups must be positive, got {groups}")
        if groups is not None:
            raise NotImplementedError("grouped matching is not built; leave groups unset to search every atom")
        if not 0.0 <= prune < 1.0:
            raise ValueError(f"prune must be in [0, 1), got {prune}")
        self.acquisition = acquisition
        self.query_chunk_size = int(query_chunk_size)
        self.dictionary_chunk_size = int(dictionary_chunk_size)
        self.top_k = int(top_k)
        self.groups = None
        self.prune = float(prune)
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
    def grouping(self) -> None:
        return None

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
        voxels = flat.shape[0]
        k = self.top_k
        indices = torch.empty((voxels, k), dtype=torch.int64, device=device)
        scores = torch.empty((voxels, k), dtype=torch.float32, device=device)
        scales = torch.empty((voxels, k), dtype=torch.complex64, device=device)
        densities = torch.empty((voxels, k), dtype=torch.float32, device=device)
        hi, lo = self._operand
        lib = _native.load()
        with torch.cuda.device(device):
            _native.check(lib.tsim_match(
                hi.data_ptr(), lo.data_ptr(), self._mode, self.dictionary.shape[0], self.dictionary.data_ptr(),
                int(torch.is_complex(self.dictionary)), self._norm.data_ptr(), flat.data_ptr(), int(signals_complex),
                voxels, contrasts, k, 0, indices.data_ptr(), scores.data_ptr(), scales.data_ptr(),
                densities.data_ptr(), _native.stream_of(device)))
        if not signals_complex and not torch.is_complex(self.dictionary):
            scales = scales.real.contiguous()
        shape = (*sample_shape, k)
        parameters = None
        if self.parameter_values.numel():
            parameters = self.parameter_values[indices].reshape(*shape, self.parameter_values.shape[-1])
        return MatchResult(parameters=parameters, indices=indices.reshape(shape), scores=scores.reshape(shape),
                           scales=scales.reshape(shape), densities=densities.reshape(shape))

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
