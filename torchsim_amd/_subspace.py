"""Temporal subspaces: a few orthonormal directions over a sequence's samples.

Interface of the reference's ``Subspace`` (src/torchsim/_subspace.py:24-129):
the same four fields, the same ``rank``/``contrasts``/``retained``/``modes``
readings, ``project`` = samples times the basis, ``expand`` = coefficients
times the basis's conjugate transpose, and ``fit`` from the left singular
vectors of a set of signals.

How it is computed here:

* a dictionary with many more atoms than samples (every MRF dictionary) is
  fitted from the samples' Gram matrix, accumulated in FP64 over blocks of
  atoms on the signals' device, then one small Hermitian eigenproblem; the
  full SVD is kept for short, wide sets;
* the projection of a simulation needs no full-length signal at all:
  ``Simulator.project(subspace, ...)`` contracts each recorded sample with the
  basis inside the real-subspace kernel.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import torch

__all__ = ["Subspace"]

#: A set of signals with at least this many atoms per sample is fitted from its
#: Gram matrix (a million-atom dictionary has a few hundred samples).
_GRAM_RATIO = 8

#: Atoms widened to FP64 at a time while the Gram matrix is summed: 65,536
#: atoms of 1,000 complex samples are 1 GB in complex128.
_GRAM_BLOCK = 65536


def _flatten(signals: Any) -> torch.Tensor:
    """Every leading axis folded into one: ``(atoms, samples)``."""
    held = torch.as_tensor(signals)
    return held.reshape(-1, held.shape[-1])


def _in_common(values: torch.Tensor, basis: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Both operands in the wider of their two dtypes, on the values' device
    (a volume is large, a basis is not, so the basis moves)."""
    common = torch.promote_types(values.dtype, basis.dtype)
    return values.to(common), basis.to(device=values.device, dtype=common)


def _gram_spectrum(rows: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Left singular vectors and singular values of ``rows.mT`` (atoms x
    samples, many more atoms) from ``G = rows.mT @ conj(rows) = U S^2 U^H``.

    G is summed in FP64 on the rows' device one block of atoms at a time, so
    the widened copy never exceeds one block. The (samples x samples)
    Hermitian eigenproblem is solved on the host: the device eigensolver
    returned NaN eigenvectors (with correct eigenvalues) for a 300 x 300
    complex128 matrix of this kind. Directions come out strongest first; each
    direction's phase is the eigensolver's."""
    wide = torch.complex128 if rows.is_complex() else torch.float64
    samples = rows.shape[1]
    gram = torch.zeros(samples, samples, dtype=wide, device=rows.device)
    for start in range(0, rows.shape[0], _GRAM_BLOCK):
        block = rows[start:start + _GRAM_BLOCK].to(wide)
        gram.addmm_(block.mT, block.conj())
    energy, vectors = torch.linalg.eigh(gram.cpu())
    energy, vectors = energy.flip(0), vectors.flip(1)
    values = energy.clamp_min(0.0).sqrt()
    return (vectors.to(device=rows.device, dtype=rows.dtype),
            values.to(device=rows.device, dtype=rows.real.dtype))


@dataclass(frozen=True)
class Subspace:
    """Orthonormal temporal directions and the spectrum they were read from.

    Attributes
    ----------
    basis : torch.Tensor
        ``(contrasts, rank)``; its columns are orthonormal, and complex when
        the fitted signals were complex.
    singular_values : torch.Tensor
        The whole spectrum of the fitted signals, strongest first (the
        directions not kept included, so another rank can be priced from it).
    dictionary : torch.Tensor, optional
        ``(atoms, contrasts)`` signals simulated to fit the basis, if any.
    simulation : Any, optional
        The record of that simulation, if any.
    """

    basis: torch.Tensor
    singular_values: torch.Tensor
    dictionary: torch.Tensor | None = None
    simulation: Any | None = None

    @property
    def rank(self) -> int:
        """Number of directions kept (columns of ``basis``)."""
        return int(self.basis.shape[-1])

    @property
    def contrasts(self) -> int:
        """Number of samples per signal the basis spans (rows of ``basis``)."""
        return int(self.basis.shape[0])

    @property
    def retained(self) -> float:
        """Share of the fitted signals' energy inside the basis: sum of the
        kept squared singular values over the sum of all of them, which is
        exactly one minus the relative squared error of a project-and-expand
        round trip of those signals (reference _subspace.py:64-73)."""
        power = self.singular_values.to(torch.float64) ** 2
        whole = float(power.sum())
        if whole == 0.0:
            return 1.0
        return float(power[: self.rank].sum()) / whole

    @property
    def modes(self) -> torch.Tensor:
        """``(rank, contrasts)``: the basis transposed, NOT conjugated -- the
        orientation subspace reconstruction operators take, which expand as
        ``sum_k conj(modes[k, t]) c_k`` (reference _subspace.py:75-84)."""
        return self.basis.transpose(-2, -1)

    def project(self, signals: torch.Tensor) -> torch.Tensor:
        """Coefficients of ``(..., contrasts)`` signals: ``signals @ basis``,
        ``(..., rank)``, in the wider dtype (reference _subspace.py:86-95)."""
        values, basis = _in_common(signals, self.basis)
        return torch.matmul(values, basis)

    def expand(self, coefficients: torch.Tensor) -> torch.Tensor:
        """Signals back from ``(..., rank)`` coefficients:
        ``coefficients @ basis^H``, ``(..., contrasts)``."""
        values, basis = _in_common(coefficients, self.basis)
        return torch.matmul(values, basis.conj().transpose(-2, -1))

    @classmethod
    def fit(cls, signals: torch.Tensor, rank: int) -> Subspace:
        """The ``rank`` strongest left singular directions of the signals
        (``(..., contrasts)``, leading axes flattened) over the contrast axis
        (reference _subspace.py:97-129). Sets of at least 8 atoms per
        contrast are fitted from their FP64 Gram matrix, which gives the same
        directions and spectrum up to each direction's phase.

        Raises
        ------
        ValueError
            If ``rank`` is below one ("rank must be positive"), or above the
            smaller dimension of the flattened signals ("exceeds").
        """
        if rank < 1:
            raise ValueError(f"rank must be positive: asked for {rank} directions")
        rows = _flatten(signals)
        atoms, samples = rows.shape
        if rank > min(atoms, samples):
            raise ValueError(f"a basis of {rank} directions exceeds what {atoms} signals of "
                             f"{samples} samples can span")
        if atoms >= _GRAM_RATIO * samples:
            vectors, spectrum = _gram_spectrum(rows)
        else:
            vectors, spectrum, _ = torch.linalg.svd(rows.transpose(-2, -1), full_matrices=False)
        return cls(basis=vectors[:, :rank], singular_values=spectrum)
