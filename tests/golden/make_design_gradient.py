"""Writes design_gradient_1024x880.npz: the sequence-design gradient of the
reference's own kernels at the full design shape (1,024 atoms x 880 pulses,
S = 10, flips U[2, 70] deg and TRs U[8, 15] ms from seed 1, T1/T2 from seed
0; loss sum tr(F^-1) of (T1, T2, M0) at sigma = 1).

The reference CPU kernel built from its sources (``make -C oracle ref``)
computes the Jacobian columns and the forward-over-reverse passes
(``oracle.reference_crlb_gradient``); the fixture holds the loss and the
gradient with respect to the 880 flips (per degree) and 880 TRs (per ms).
Run from the repository root: ``python tests/golden/make_design_gradient.py``.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from tests import _cases  # noqa: E402


def design_inputs(pulses: int, atoms: int, seed: int = 1):
    rng = np.random.default_rng(seed)
    flips = rng.uniform(2.0, 70.0, pulses)
    trs = rng.uniform(8.0, 15.0, pulses)
    t1, t2 = _cases.random_tissue(atoms, seed=0)
    return flips, trs, t1, t2


def reference_gradient(flips, trs, t1, t2, states=10, threads=8):
    from torchsim_amd.simulators import MRFSimulator

    packed = MRFSimulator().packed(flip=torch.tensor(flips, dtype=torch.float32),
                                   TR=torch.tensor(trs, dtype=torch.float32))
    events = oracle.events_of(packed)
    loss, g_flip, g_duration = oracle.reference_crlb_gradient(events, t1, t2, states=states, outputs=len(flips),
                                                              threads=threads)
    rf = np.flatnonzero((events["kind"] == 1) & ((events["action"] & 4) == 0))
    per_degree = g_flip[rf] * np.pi / 180.0
    per_ms = np.zeros(len(trs))
    per_ms[:-1] = g_duration[rf[1:]] * 1e-3  # the pulse after readout k carries TR_k
    return loss, per_degree, per_ms


def main() -> None:
    flips, trs, t1, t2 = design_inputs(880, 1024)
    loss, g_flip, g_tr = reference_gradient(flips, trs, t1, t2)
    out = Path(__file__).resolve().parent / "design_gradient_1024x880.npz"
    np.savez_compressed(out, loss=np.float64(loss), grad_flip=g_flip, grad_tr=g_tr)
    print(f"wrote {out}: loss {loss:.9e}")


if __name__ == "__main__":
    main()
