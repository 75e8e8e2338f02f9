"""Isochromat (Bloch) brute force -- TEST INFRASTRUCTURE ONLY.

An independent model of the same experiment, used to pin the EPG oracle: a
voxel is N isochromats spread uniformly over one cycle of dephasing; each
carries a real magnetization vector (Mx, My, Mz). Relaxation is the Bloch
solution over the interval, a pulse is a rigid rotation of the vector, an
unbalanced gradient advances isochromat j's transverse phase by 2 pi j / N,
spoiling zeroes the transverse part, an inversion scales Mz, and a sample is
the average of Mx + i My demodulated by the sample phase.

EPG with enough orders that nothing is truncated equals this average exactly
as long as N exceeds twice the highest order reached, so agreement pins the
EPG shift direction, the T-matrix, and the relaxation/recovery terms without
sharing a line of code with them.
"""

from __future__ import annotations

import numpy as np

_PRE_SHIFT, _POST_SHIFT, _INVERSION, _SPOIL_AFTER, _SHIFT_AFTER, _RECORD = 1, 2, 4, 8, 16, 32


def _rotation(alpha: float, phi: float) -> np.ndarray:
    """Rotation by alpha about the transverse axis at angle phi (right-handed
    about n = (cos phi, sin phi, 0)), as a 3x3 matrix on (Mx, My, Mz)."""
    n = np.array([np.cos(phi), np.sin(phi), 0.0])
    k = np.array([[0.0, -n[2], n[1]], [n[2], 0.0, -n[0]], [-n[1], n[0], 0.0]])
    theta = alpha
    return np.eye(3) + np.sin(theta) * k + (1.0 - np.cos(theta)) * (k @ k)


def simulate(
    events: dict,
    *,
    t1: float,
    t2: float,
    m0: float = 1.0,
    b1: float = 1.0,
    ie: float = 1.0,
    b0: float = 0.0,
    isochromats: int = 512,
) -> np.ndarray:
    """Complex samples (outputs,) for one tissue."""
    n = isochromats
    theta = 2.0 * np.pi * np.arange(n) / n
    m = np.zeros((n, 3))
    m[:, 2] = 1.0
    outputs = int(np.max(events["output_index"])) + 1 if np.any(events["output_index"] >= 0) else 0
    samples = np.zeros(outputs, dtype=np.complex128)
    phases = events.get("phase", np.zeros(len(events["kind"])))

    def shift():
        c, s = np.cos(theta), np.sin(theta)
        x, y = m[:, 0].copy(), m[:, 1].copy()
        m[:, 0] = c * x - s * y
        m[:, 1] = s * x + c * y

    for e in range(len(events["kind"])):
        dt = float(events["duration"][e])
        e1 = np.exp(-1000.0 / t1 * dt)
        e2 = np.exp(-1000.0 / t2 * dt)
        m[:, 0:2] *= e2
        m[:, 2] = m[:, 2] * e1 + (1.0 - e1)
        if b0:
            # free precession in the rotating frame: Mxy turns by -2 pi b0 dt
            c, s = np.cos(-2 * np.pi * b0 * dt), np.sin(-2 * np.pi * b0 * dt)
            x, y = m[:, 0].copy(), m[:, 1].copy()
            m[:, 0] = c * x - s * y
            m[:, 1] = s * x + c * y
        act = int(events["action"][e])
        kind = int(events["kind"][e])
        if act & _PRE_SHIFT:
            shift()
        if kind == 1:
            if act & _INVERSION:
                m[:, 2] *= -ie
            else:
                rot = _rotation(float(events["flip"][e]) * b1, float(phases[e]))
                m[:] = m @ rot.T
        elif kind == 2 and act & _RECORD:
            transverse = (m[:, 0] + 1j * m[:, 1]).mean()
            samples[int(events["output_index"][e])] = m0 * transverse * np.exp(-1j * float(phases[e]))
        if act & _POST_SHIFT:
            shift()
        if act & _SPOIL_AFTER:
            m[:, 0:2] = 0.0
        elif act & _SHIFT_AFTER:
            shift()
    return samples
