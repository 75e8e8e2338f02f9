"""FP64 temporal basis and projection -- TEST INFRASTRUCTURE ONLY.

Restates the reference's ``Subspace`` (src/torchsim/_subspace.py:24-129) in
numpy float64 / complex128:

* ``fit``: the leading ``rank`` left singular vectors of the flattened
  signals' transpose (``svd(flat.T)``, :97-129), and every singular value;
* ``project``: ``signals @ basis`` (a plain transpose, :86-95);
* ``expand``: ``coefficients @ basis^H`` (:97-99);
* ``retained``: kept energy over total energy of the singular values (:64-73).

Pinned by the reference's own tests (tests/estimators/test_subspace.py:
one minus ``retained`` is the measured relative squared error of the round
trip at every rank; a full-rank basis round-trips exactly; the basis is
orthonormal; an echo train needs very few directions), and by the round
trip's independence of each direction's phase, which a planted conjugated
projection fails (as does a retained fraction of unsquared values).
"""

from __future__ import annotations

import numpy as np


def fit(signals: np.ndarray, rank: int) -> tuple[np.ndarray, np.ndarray]:
    """(basis (contrasts, rank), singular values) of ``signals (..., contrasts)``."""
    if rank < 1:
        raise ValueError(f"rank must be positive, got {rank}")
    flat = np.asarray(signals).reshape(-1, np.asarray(signals).shape[-1])
    if rank > min(flat.shape):
        raise ValueError(f"rank={rank} exceeds the signals' dimensions {flat.shape}")
    wide = flat.astype(np.complex128) if np.iscomplexobj(flat) else flat.astype(np.float64)
    u, s, _ = np.linalg.svd(wide.T, full_matrices=False)
    return u[:, :rank], s


def project(signals: np.ndarray, basis: np.ndarray) -> np.ndarray:
    """``signals @ basis`` in the wider dtype."""
    return np.asarray(signals) @ basis


def expand(coefficients: np.ndarray, basis: np.ndarray) -> np.ndarray:
    """``coefficients @ basis^H``."""
    return np.asarray(coefficients) @ basis.conj().T


def retained(singular_values: np.ndarray, rank: int) -> float:
    """The fraction of the fitted energy the first ``rank`` directions keep."""
    energy = np.asarray(singular_values, dtype=np.float64) ** 2
    total = float(energy.sum())
    return 1.0 if total == 0.0 else float(energy[:rank].sum()) / total
