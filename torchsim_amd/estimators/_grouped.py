"""Grouped dictionary matching (reference estimators/_grouped.py:33-245).

The dictionary is split into groups of nearly parallel atoms, each with one
representative (the normalized mean of its normalized atoms); a voxel is
scored against the representatives first and searched only inside the
groups whose representative scores within ``prune`` of its best (Cauley
et al., "Fast group matching for MR fingerprinting reconstruction", MRM 74,
2015). The grouping is the reference's greedy one -- repeatedly take the
first unassigned atom and give it the ``ceil(atoms / count)`` unassigned
atoms most parallel to it -- run on the device: one matrix-vector product
over the dictionary per group, assigned atoms masked out, instead of
gathering the unassigned ones again for every group. Inside a group the
search is the fused MFMA matcher (``DictionaryMatcher`` on the group's own
atoms), whose candidates are scored exactly in FP64.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

__all__ = ["Grouping", "correlate"]


def _unit_rows(values: torch.Tensor) -> torch.Tensor:
    """Rows scaled to unit norm (the norm floored at float32 epsilon, so an
    all-zero row stays zero)."""
    norm = torch.linalg.vector_norm(values, dim=-1, keepdim=True)
    return values / norm.clamp_min(torch.finfo(torch.float32).eps)


def correlate(signals: torch.Tensor, atoms: torch.Tensor) -> torch.Tensor:
    """|<s, d>| of every (normalized) signal against every (normalized) atom:
    ``(voxels, atoms)``, real. A complex signal against real atoms correlates
    its two parts separately, so the score does not depend on its phase."""
    if torch.is_complex(signals) and not torch.is_complex(atoms):
        return torch.hypot(signals.real @ atoms.mT, signals.imag @ atoms.mT)
    if torch.is_complex(atoms) and not torch.is_complex(signals):
        signals = signals.to(atoms.dtype)
    return (signals @ atoms.mH).abs()


@dataclass(frozen=True)
class Grouping:
    """A dictionary partitioned into groups of nearly parallel atoms.

    Attributes
    ----------
    members : tuple of torch.Tensor
        The atom indices of each group.
    representative : torch.Tensor
        ``(groups, contrasts)``: each group's normalized mean signal.
    atoms : tuple of torch.Tensor
        Each group's atoms, normalized, in the basis the dictionary came in.
    """

    members: tuple[torch.Tensor, ...]
    representative: torch.Tensor
    atoms: tuple[torch.Tensor, ...]

    @property
    def count(self) -> int:
        """Number of groups."""
        return len(self.members)

    @property
    def sizes(self) -> tuple[int, ...]:
        """Atoms per group."""
        return tuple(int(index.numel()) for index in self.members)

    @property
    def condition(self) -> float:
        """Largest over smallest singular value of the representatives: it
        climbs when they start to resemble one another (too many groups)."""
        values = torch.linalg.svdvals(self.representative.to(torch.complex64))
        return float(values[0] / values[-1].clamp_min(torch.finfo(torch.float32).tiny))

    @classmethod
    def fit(cls, dictionary: torch.Tensor, count: int) -> Grouping:
        """Partition ``(atoms, contrasts)`` into ``count`` groups, greedily
        (every group but the last holds ``ceil(atoms / count)`` atoms).

        Raises
        ------
        ValueError
            If ``count`` is not between one and the number of atoms.
        """
        total = int(dictionary.shape[0])
        if not 1 <= count <= total:
            raise ValueError(f"count must be between one and the {total} atoms, got {count}")
        unit = _unit_rows(dictionary)
        share = -(-total // count)
        free = torch.ones(total, dtype=torch.bool, device=dictionary.device)
        members = []
        left = total
        # the first free atom of each round, read back in a few batched waits
        while left:
            seed = int(torch.nonzero(free, as_tuple=False)[0, 0])
            alike = (unit @ unit[seed].conj()).abs()
            alike = torch.where(free, alike, torch.full_like(alike, -1.0))
            take = min(share, left)
            chosen = torch.topk(alike, take).indices
            members.append(chosen)
            free[chosen] = False
            left -= take
        representative = torch.stack([_unit_rows(unit[index].mean(0, keepdim=True))[0] for index in members])
        return cls(members=tuple(members), representative=representative,
                   atoms=tuple(unit[index] for index in members))

    def survivors(self, signals: torch.Tensor, prune: float) -> torch.Tensor:
        """``(voxels, groups)`` bool: the groups whose representative scores
        at least ``(1 - prune)`` of the voxel's best (signals normalized)."""
        scores = correlate(signals, self.representative)
        best = scores.amax(-1, keepdim=True)
        return scores >= best * (1.0 - prune)

    def to(self, device: torch.device) -> Grouping:
        """This grouping on ``device``."""
        if self.representative.device == torch.device(device):
            return self
        return Grouping(members=tuple(index.to(device) for index in self.members),
                        representative=self.representative.to(device),
                        atoms=tuple(value.to(device) for value in self.atoms))
