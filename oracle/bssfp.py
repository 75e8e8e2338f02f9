"""Balanced SSFP steady state by a fixed-point solve -- TEST INFRASTRUCTURE ONLY.

Checks the closed form of ``bSSFPSimulator`` (reference
src/torchsim/simulators/bssfp.py:74-121) from first principles, sharing no
formula with it: one repetition maps the magnetization vector M (just after
the pulse) to R_x(flip) (E P(theta) M + (1 - E1) M0 z), where P(theta) turns
the transverse plane by theta = 2 pi (chemshift - B0) TR + phase_inc,
E = diag(E2, E2, E1) with E2 = exp(-TR/T2), and R_x turns y towards z. The
steady state solves (I - R_x E P) M = R_x (1 - E1) M0 z; the echo a time TE
later is (Mx + i My) exp(-TE/T2) exp(i 2 pi (chemshift - B0) TE).

The reference's closed form takes E2 at TE inside the steady state as well,
so the two agree exactly when TE = TR; that is where this pins it (a dropped
term, a wrong sign or a swapped cosine fails by far more than rounding).
Elsewhere the closed form is pinned by restatement only.
"""

from __future__ import annotations

import numpy as np


def _turn(theta: float) -> np.ndarray:
    """Free precession by theta: x turns towards y."""
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def steady_state(t1: float, t2: float, m0: float, b0: float, chemshift: float, flip_deg: float, tr_ms: float,
                 te_ms: float, phase_inc_deg: float) -> complex:
    """The echo of the balanced steady state (units as the simulator's)."""
    alpha = np.deg2rad(flip_deg)
    omega = 2.0 * np.pi * (chemshift - b0)
    theta = omega * tr_ms * 1e-3 + np.deg2rad(phase_inc_deg)
    e1, e2 = np.exp(-tr_ms / t1), np.exp(-tr_ms / t2)
    c, s = np.cos(alpha), np.sin(alpha)
    rf = np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])
    turn = _turn(theta)
    decay = np.diag([e2, e2, e1])
    m = np.linalg.solve(np.eye(3) - rf @ decay @ turn, rf @ np.array([0.0, 0.0, m0 * (1.0 - e1)]))
    return complex(m[0], m[1]) * np.exp(-te_ms / t2) * np.exp(1j * omega * te_ms * 1e-3)
