"""A low-rank temporal basis of simulated signals.

The reference's ``Subspace`` (src/torchsim/_subspace.py:24-129): the leading
left singular vectors of a set of signals, the singular values they came with,
and the two maps between contrasts and coefficients. Projecting is
``signals @ basis`` and expanding ``coefficients @ basis^H`` (a plain, not a
conjugate, transpose on the way in, as the reference and the reconstruction
libraries that read ``modes`` define it).

On this library the projection also runs fused into the real-subspace
kernel: ``Simulator.project(subspace, ...)`` contracts each recorded sample
with the basis as the kernel produces it, so a dictionary or a Jacobian is
never written out at full length.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import torch

__all__ = ["Subspace"]


@dataclass(frozen=True)
class Subspace:
    """A temporal basis fitted to a set of signals.

    Attributes
    ----------
    basis : torch.Tensor
        ``(contrasts, rank)``, orthonormal; complex when the signals are.
    singular_values : torch.Tensor
        Every singular value of the fitted signals, not only the kept ones.
    dictionary : torch.Tensor, optional
        The signals the basis was fitted to, when simulated for the purpose.
    simulation : Any, optional
        Whatever recorded that simulation.
    """

    basis: torch.Tensor
    singular_values: torch.Tensor
    dictionary: torch.Tensor | None = None
    simulation: Any | None = None

    @property
    def rank(self) -> int:
        """How many directions the basis keeps."""
        return int(self.basis.shape[-1])

    @property
    def contrasts(self) -> int:
        """How many contrasts it was fitted over."""
        return int(self.basis.shape[0])

    @property
    def retained(self) -> float:
        """The fraction of the fitted signals' energy the basis keeps: one minus
        the relative squared error of projecting them onto it and back
        (reference _subspace.py:64-73)."""
        energy = self.singular_values.to(torch.float64).square()
        total = float(energy.sum())
        return 1.0 if total == 0.0 else float(energy[: self.rank].sum()) / total

    @property
    def modes(self) -> torch.Tensor:
        """The basis rank-first, ``(rank, contrasts)``: a plain transpose
        (reference _subspace.py:75-84)."""
        return self.basis.mT

    def project(self, signals: torch.Tensor) -> torch.Tensor:
        """``(..., contrasts)`` to ``(..., rank)``: ``signals @ basis`` in the
        wider of the two dtypes, on the signals' device (reference
        _subspace.py:86-95)."""
        dtype = torch.promote_types(signals.dtype, self.basis.dtype)
        return signals.to(dtype) @ self.basis.to(dtype).to(signals.device)

    def expand(self, coefficients: torch.Tensor) -> torch.Tensor:
        """Coefficients back to contrasts: ``coefficients @ basis^H``."""
        dtype = torch.promote_types(coefficients.dtype, self.basis.dtype)
        return coefficients.to(dtype) @ self.basis.mH.to(dtype).to(coefficients.device)

    @classmethod
    def fit(cls, signals: torch.Tensor, rank: int) -> Subspace:
        """The leading ``rank`` left singular vectors of the signals'
        transpose; every leading axis is flattened (reference
        _subspace.py:97-129). With at least 8 atoms per contrast they come
        from the contrasts' Gram matrix in double precision (the same
        directions and singular values, up to each direction's phase)
        rather than from an SVD of the whole dictionary.

        Raises
        ------
        ValueError
            If ``rank`` is not positive or exceeds what the signals span.
        """
        if rank < 1:
            raise ValueError(f"rank must be positive, got {rank}")
        signals = torch.as_tensor(signals)
        flat = signals.reshape(-1, signals.shape[-1])
        if rank > min(flat.shape):
            raise ValueError(f"rank={rank} exceeds the signals' dimensions {tuple(flat.shape)}")
        if flat.shape[0] >= _GRAM_RATIO * flat.shape[1]:
            basis, singular_values = _tall(flat)
        else:
            basis, singular_values, _ = torch.linalg.svd(flat.mT, full_matrices=False)
        return cls(basis=basis[:, :rank], singular_values=singular_values)


#: Atoms per contrast beyond which the basis comes from the contrasts' Gram
#: matrix: a dictionary of a million atoms has a few hundred contrasts.
_GRAM_RATIO = 8


def _tall(flat: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Left singular vectors and singular values of ``flat.mT`` for many more
    rows (atoms) than columns (contrasts), from the eigen-decomposition of the
    (contrasts x contrasts) matrix ``flat.mT @ conj(flat)`` = U S^2 U^H, formed
    in double precision on the signals' device (one GEMM over the atoms
    instead of an SVD of the whole dictionary). The small Hermitian
    eigenproblem is solved on the host: the device eigensolver returned NaN
    eigenvectors for such a matrix (complex128, 300 x 300) with correct
    eigenvalues. Directions come out in decreasing order; each direction's
    phase is the eigensolver's."""
    wide = torch.complex128 if flat.is_complex() else torch.float64
    work = flat.to(wide)
    gram = (work.mT @ work.conj()).cpu()
    energy, vectors = torch.linalg.eigh(gram)
    energy, vectors = energy.flip(0), vectors.flip(1)
    return (vectors.to(device=flat.device, dtype=flat.dtype),
            energy.clamp_min(0.0).sqrt().to(device=flat.device, dtype=flat.real.dtype))
