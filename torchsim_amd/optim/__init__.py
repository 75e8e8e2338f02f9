"""Sequence design: the Cramer-Rao bound and a bounded optimizer over sequence
parameters (reference optim/_design.py)."""

from ._design import Bounded, SequenceDesign, SequenceOptimization, crlb

__all__ = ["Bounded", "SequenceDesign", "SequenceOptimization", "crlb"]
