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


def grouping(dictionary, count: int):
    """The reference's greedy grouping (estimators/_grouped.py:94-134), FP64:
    normalize the atoms; while atoms are unassigned, take the first of them
    and give it the ceil(atoms / count) unassigned atoms with the largest
    |<a, seed>| (the seed itself first). Returns (members, representative),
    members a list of index arrays and representative the normalized mean of
    each group's normalized atoms."""
    d = np.asarray(dictionary)
    d = d.astype(np.complex128) if np.iscomplexobj(d) else d.astype(np.float64)
    unit = d / np.maximum(np.linalg.norm(d, axis=1), EPS32)[:, None]
    total = d.shape[0]
    share = -(-total // count)
    unassigned = np.arange(total)
    members = []
    while unassigned.size:
        pool = unit[unassigned]
        alike = np.abs(pool @ np.conj(pool[0]))
        take = min(share, unassigned.size)
        chosen = np.argsort(-alike, kind="stable")[:take]
        members.append(unassigned[chosen])
        unassigned = np.delete(unassigned, chosen)
    means = np.stack([unit[m].mean(axis=0) for m in members])
    representative = means / np.maximum(np.linalg.norm(means, axis=1), EPS32)[:, None]
    return members, representative


def survivors(representative, signals, prune: float):
    """(voxels, groups) bool: |<y, r>| / |y| >= (1 - prune) max over groups."""
    y = np.asarray(signals)
    y = y.astype(np.complex128) if np.iscomplexobj(y) else y.astype(np.float64)
    unit = y / np.maximum(np.linalg.norm(y, axis=1), EPS32)[:, None]
    score = np.abs(unit @ np.conj(np.asarray(representative)).T)
    return score >= score.max(axis=1, keepdims=True) * (1.0 - prune)


def match_grouped(dictionary, signals, count: int, prune: float, top_k: int = 1):
    """The grouped search of the reference (estimators/_grouped.py:164-215):
    each voxel scored exhaustively (as :func:`match`) against the atoms of
    its surviving groups only. Returns (indices, scores), (voxels, top_k);
    a voxel with fewer surviving atoms than top_k keeps (0, -1) padding."""
    d = np.asarray(dictionary)
    y = np.asarray(signals).reshape(-1, d.shape[1])
    members, representative = grouping(d, count)
    keep = survivors(representative, y, prune)
    indices = np.zeros((y.shape[0], top_k), dtype=np.int64)
    scores = np.full((y.shape[0], top_k), -1.0)
    for v in range(y.shape[0]):
        allowed = np.concatenate([members[g] for g in np.nonzero(keep[v])[0]])
        idx, sc, _, _ = match(d[allowed], y[v:v + 1], top_k=min(top_k, allowed.size))
        found = allowed[idx[0]]
        # the lower atom index first among equal scores, as match() ranks
        order = np.lexsort((found, -sc[0]))
        indices[v, : found.size] = found[order]
        scores[v, : found.size] = sc[0][order]
    return indices, scores
