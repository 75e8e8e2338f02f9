"""The exported functionals keep the reference's parameter lists (so a
positional call lands where it would there), and the bSSFP closed form behind
``bssfp_sim`` agrees with a first-principles FP64 fixed-point solve. CPU only."""

from __future__ import annotations

import inspect

import numpy as np
import pytest
import torch

from oracle import bssfp as oracle_bssfp
from torchsim_amd import bssfp_sim, fse_sim, mrf_sim
from torchsim_amd.simulators.bssfp import steady_state

# (name, default) in order; inspect.Parameter.empty for none. Literals read
# off reference _functional/_mrf.py:12-25, _bssfp.py:12-24, _fse.py:12-26.
_NONE = inspect.Parameter.empty
SIGNATURES = {
    mrf_sim: [("flip", _NONE), ("TR", _NONE), ("T1", _NONE), ("T2", _NONE), ("diff", None), ("B1", 1.0),
              ("inv_efficiency", 1.0), ("M0", 1.0), ("TI", 0.0), ("states", 10), ("nreps", 1), ("device", None)],
    bssfp_sim: [("flip", _NONE), ("TE", _NONE), ("TR", _NONE), ("T1", _NONE), ("T2", _NONE), ("diff", None),
                ("phase_inc", 180.0), ("B0", 0.0), ("chemshift", 0.0), ("M0", 1.0), ("device", None)],
    fse_sim: [("flip", _NONE), ("ESP", _NONE), ("T1", _NONE), ("T2", _NONE), ("phases", 0.0), ("TR", 1e6),
              ("exc_flip", 90.0), ("exc_phase", 90.0), ("diff", None), ("B1", 1.0), ("M0", 1.0), ("states", 10),
              ("device", None)],
}


@pytest.mark.parametrize("function", list(SIGNATURES), ids=lambda f: f.__name__)
def test_parameter_lists_are_the_references(function):
    got = [(p.name, p.default) for p in inspect.signature(function).parameters.values()]
    assert got == SIGNATURES[function]


def _cases(count, seed):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        yield dict(t1=rng.uniform(200, 3000), t2=rng.uniform(10, 300), m0=rng.uniform(0.5, 1.5),
                   b0=rng.uniform(-80, 80), chemshift=rng.choice([0.0, -220.0]), flip_deg=rng.uniform(2, 90),
                   tr_ms=rng.uniform(2, 12), phase_inc_deg=rng.choice([0.0, 180.0, rng.uniform(0, 360)]))


def _closed(case, te):
    t = {k: torch.tensor(float(v), dtype=torch.float64) for k, v in case.items()}
    return complex(steady_state(t["t1"], t["t2"], t["m0"], t["b0"], t["chemshift"], t["flip_deg"], t["tr_ms"],
                                torch.tensor(te, dtype=torch.float64), t["phase_inc_deg"]))


def test_closed_form_is_the_fixed_point_of_one_repetition():
    """At TE = TR the reference's closed form is the steady state of the
    repetition map solved as a 3x3 linear system (oracle/bssfp.py)."""
    for case in _cases(200, seed=5):
        ref = oracle_bssfp.steady_state(**case, te_ms=case["tr_ms"])
        got = _closed(case, case["tr_ms"])
        assert abs(got - ref) <= 1e-12 * max(abs(ref), 1e-3), (case, got, ref)


def test_a_planted_mistake_in_the_fixed_point_fails(monkeypatch):
    """The comparison above can fail: with the oracle's precession turned the
    other way, off-resonant cases disagree."""
    turn = oracle_bssfp._turn
    monkeypatch.setattr(oracle_bssfp, "_turn", lambda theta: turn(-theta))
    worst = max(abs(_closed(c, c["tr_ms"]) - oracle_bssfp.steady_state(**c, te_ms=c["tr_ms"])) for c in _cases(50, 6))
    assert worst > 1e-3


def test_on_resonance_alternating_magnitude():
    """On resonance with 180 degree cycling and TE = TR the steady state is
    the textbook M0 (1 - E1) sin a / (1 - (E1 - E2) cos a - E1 E2) times E2."""
    t1, t2, tr, flip = 1000.0, 100.0, 5.0, 40.0
    e1, e2, a = np.exp(-tr / t1), np.exp(-tr / t2), np.deg2rad(flip)
    textbook = (1 - e1) * np.sin(a) / (1 - (e1 - e2) * np.cos(a) - e1 * e2) * e2
    got = _closed(dict(t1=t1, t2=t2, m0=1.0, b0=0.0, chemshift=0.0, flip_deg=flip, tr_ms=tr, phase_inc_deg=180.0), tr)
    assert abs(got.real) < 1e-15 and got.imag == pytest.approx(textbook, rel=1e-13)
