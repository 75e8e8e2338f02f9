"""FP64 exhaustive dictionary matcher -- TEST INFRASTRUCTURE ONLY.

Restates the reference's matching (estimators/_dictionary.py:333-431 and
estimators/_grouped.py:219-245): the score of voxel y against atom d is
|<y, d>| / (|y| |d|) with both norms floored at float32 epsilon, a real atom
scoring a complex voxel by hypot(<Re y, d>, <Im y, d>) (= |<y, d>|); the
best top_k atoms by score (ties: the lower atom index first, an order the
reference leaves to torch.topk); the complex least-squares scale
<d, y> / |d|^2 and the density |y| score / |d| of each.

Pinned by the reference's own known answers (tests/estimators/
test_dictionary.py:12-52: complex-scaled atoms recovered with score 1 and
their exact scale; an identity dictionary), by phase invariance of the score
(test_dictionary.py:55-99, which a real-part-only score fails), and by the
least-squares property (the residual y - scale d is orthogonal to d).
"""

from __future__ import annotations

import numpy as np

EPS32 = float(np.finfo(np.float32).eps)


def _inner(d, y):
    """<d, y> = sum conj(d) y for every (voxel, atom), and the floored norms.
    Equal atoms get bitwise equal columns (computed once per distinct row),
    so an exact tie stays a tie."""
    distinct, back = np.unique(d, axis=0, return_inverse=True)
    back = np.asarray(back).reshape(-1)
    inner = (y @ np.conj(distinct).T)[:, back]
    dn = np.maximum(np.linalg.norm(distinct, axis=1), EPS32)[back]
    yn = np.maximum(np.linalg.norm(y, axis=1), EPS32)
    return inner, dn, yn


def match(dictionary, signals, top_k: int = 1):
    """(indices, scores, scales, densities), each (voxels, top_k), FP64."""
    d = np.asarray(dictionary)
    y = np.asarray(signals)
    d = d.astype(np.complex128) if np.iscomplexobj(d) else d.astype(np.float64)
    y = y.astype(np.complex128) if np.iscomplexobj(y) else y.astype(np.float64)
    y = y.reshape(-1, d.shape[1])
    inner, dn, yn = _inner(d, y)
    score = np.abs(inner) / (yn[:, None] * dn[None, :])
    atoms = np.arange(d.shape[0])
    indices = np.empty((y.shape[0], top_k), dtype=np.int64)
    for v in range(y.shape[0]):
        order = np.lexsort((atoms, -score[v]))
        indices[v] = order[:top_k]
    rows = np.arange(y.shape[0])[:, None]
    scores = score[rows, indices]
    scales = inner[rows, indices] / dn[indices] ** 2
    densities = yn[:, None] * scores / dn[indices]
    return indices, scores, scales, densities


def gaps(dictionary, signals, top_k: int = 1):
    """Per voxel, the smallest score gap between consecutive ranks among the
    first top_k + 1 (how far the top_k ranking is from a tie); +inf when the
    dictionary has no more than top_k atoms."""
    d = np.asarray(dictionary)
    y = np.asarray(signals).reshape(-1, d.shape[1])
    d = d.astype(np.complex128) if np.iscomplexobj(d) else d.astype(np.float64)
    y = y.astype(np.complex128) if np.iscomplexobj(y) else y.astype(np.float64)
    inner, dn, yn = _inner(d, y)
    score = np.abs(inner) / (yn[:, None] * dn[None, :])
    if d.shape[0] <= top_k:
        return np.full(y.shape[0], np.inf)
    ranked = -np.sort(-score, axis=1)[:, : top_k + 1]
    return np.min(ranked[:, :-1] - ranked[:, 1:], axis=1)
