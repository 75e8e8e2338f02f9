"""The complex EPG state machine in torch float64 -- TEST INFRASTRUCTURE ONLY.

A restatement of the reference's packed state machine
(tests/utils/packed_reference.py:36-389, src/torchsim/_epg_cpu.cpp:4216-4566)
in double precision, written as whole-array torch operations per event so
autograd (forward and reverse mode, any order) differentiates it. It carries
F+, F- and Z per order, an optional chemically exchanging second pool (its
interval operators reached through ``torch.linalg.matrix_exp`` and a linear
solve -- a different algorithm from the closed form the kernels evaluate),
per-order diffusion damping, off-resonance, and RF phase. Shifts truncate
exactly as the kernels do (top order dropped, F-[S-1] = 0).

Parity: pinned (tests/test_oracle_packed.py) -- against the C oracle on
real-subspace inputs (which the benchmark checksums pin); against the
isochromat brute force with phases and off-resonance; against the reference's
own FP32 complex kernel built from source (``oracle.reference_complex``) for
the exchanging pool with relaxation, chemical shift, B0, B1 and phases, and
for diffusion on the real and complex paths, signal and every tissue column;
by the exchange closed forms (equilibrium populations are stationary; fast
exchange coalesces at the population-weighted frequency); and by central
differences. A planted swap of the exchange rates, or the transverse
diffusion weight planted on Z, fails those tests.
"""

from __future__ import annotations

from typing import Any

import torch

PRE_SHIFT, POST_SHIFT, INVERSION, SPOIL_AFTER, SHIFT_AFTER, RECORD = 1, 2, 4, 8, 16, 32
_C = torch.complex128
_R = torch.float64


def _shift(fp: torch.Tensor, fm: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    zero = torch.zeros_like(fm[:, :1])
    down = torch.cat((fm[:, 1:], zero), dim=1)
    up = torch.cat((down[:, :1].conj(), fp[:, :-1]), dim=1)
    return up, down


def _at0(like: torch.Tensor, value: torch.Tensor) -> torch.Tensor:
    out = torch.zeros_like(like)
    return torch.cat((value.reshape(-1, 1).to(like.dtype), out[:, 1:]), dim=1)


def simulate(
    events: dict[str, Any],
    tissue: dict[str, Any],
    *,
    states: int,
    outputs: int,
    lineshape: tuple[Any, Any, float] | None = None,
    profile: tuple[Any, Any, Any, Any, float] | None = None,
    _plant: str | None = None,
) -> torch.Tensor:
    """Recorded complex samples (atoms, outputs), complex128.

    ``events``: duration (s), kind, flip (rad), phase (rad), action,
    output_index -- numpy or torch, one train. ``tissue``: t1, t2 (ms) and
    optionally m0, b1, b0 (Hz), b1_phase (rad), ie, diffusion (the rate
    already scaled to 1/s per unit order^2), and for an exchanging pool
    fraction, exchange (1/s), t1b, t2b (ms), shift (Hz); for a semisolid
    (magnetization-transfer) pool bound_fraction, bound_exchange (1/s) and
    t1_bound (ms), with the events' ``saturation`` (1/s per rad^2) and
    ``rf_frequency`` (Hz) and ``lineshape`` = (values, slopes, step_hz), the
    absorption table read by cubic Hermite (reference
    src/torchsim/_epg_cpu.cpp:875-932, 1888-1910, 4216-4566 with pool kind 1:
    the bound pool has Z states only, exchanging with the free water's Z and
    saturated by exp(saturation theta^2 G(rf_frequency - b0)) at every
    pulse; an inversion leaves it alone). With ``profile`` = (a, b, slope_a,
    slope_b, theta_max), each (locations, bins) complex, every pulse but an
    inversion turns the states through the Cayley-Klein pair read by cubic
    Hermite at the atom's position (atom % locations) and effective flip,
    b turned by exp(-i phi) (reference src/torchsim/_epg_cpu.cpp:2041-2049,
    2107-2160, 2273-2304). Tissue entries may be tensors requiring grad;
    flip and duration may be too.

    ``_plant`` exists only so the tests can show that their checks fail on a
    plausible mistake: ``"swap_exchange"`` exchanges the two rate constants
    (k_ab = k (1 - f), k_ba = k f), ``"z_transverse_weight"`` damps the
    longitudinal states with the transverse weight l^2 + l + 1/3.
    """
    def tt(name, default=None):
        value = tissue.get(name, default)
        if value is None:
            return None
        return torch.as_tensor(value, dtype=_R).reshape(-1)

    t1, t2 = tt("t1"), tt("t2")
    atoms = max(int(v.numel()) for v in (tt(k) for k in tissue) if v is not None)
    ones = torch.ones(atoms, dtype=_R)
    t1, t2 = t1 * ones, t2 * ones
    m0 = tt("m0", 1.0) * ones
    b1 = tt("b1", 1.0) * ones
    b0 = tt("b0", 0.0) * ones
    b1_phase = tt("b1_phase", 0.0) * ones
    ie = tt("ie", 1.0) * ones
    rate = tt("diffusion", 0.0) * ones
    exchanging = "fraction" in tissue and tissue["fraction"] is not None
    semisolid = tissue.get("bound_fraction") is not None
    duration = torch.as_tensor(events["duration"], dtype=_R)
    flip = torch.as_tensor(events["flip"], dtype=_R)
    phase = torch.as_tensor(events["phase"], dtype=_R)
    kind = [int(k) for k in torch.as_tensor(events["kind"]).tolist()]
    action = [int(a) for a in torch.as_tensor(events["action"]).tolist()]
    out_index = [int(o) for o in torch.as_tensor(events["output_index"]).tolist()]

    order = torch.arange(states, dtype=_R)
    wt = order**2 + order + 1.0 / 3.0
    wl = wt if _plant == "z_transverse_weight" else order**2
    shape = (atoms, states)
    pools = 2 if exchanging else 1
    fp = [torch.zeros(shape, dtype=_C) for _ in range(pools)]
    fm = [torch.zeros(shape, dtype=_C) for _ in range(pools)]
    z = [torch.zeros(shape, dtype=_C) for _ in range(pools)]
    if exchanging:
        f = tt("fraction") * ones
        k = tt("exchange") * ones
        r1b = 1000.0 / (tt("t1b") * ones)
        r2b = 1000.0 / (tt("t2b") * ones)
        dfb = tt("shift", 0.0) * ones
        free = 1.0 - f
        kab, kba = (k * free, k * f) if _plant == "swap_exchange" else (k * f, k * free)
        r1a, r2a = 1000.0 / t1, 1000.0 / t2
        gl = torch.stack((torch.stack((-kab - r1a, kba), -1), torch.stack((kab, -kba - r1b), -1)), -2)
        gt = torch.stack(
            (torch.stack((-kab - r2a, kba), -1), torch.stack((kab, -kba - r2b), -1)), -2
        ).to(_C) + torch.diag_embed(torch.stack((torch.zeros_like(dfb), dfb), -1) * (-2j * torch.pi))
        source = torch.stack((free * r1a, f * r1b), -1)
        settled = torch.linalg.solve(gl, source.unsqueeze(-1)).squeeze(-1)
        z[0] = _at0(z[0], free)
        z[1] = _at0(z[1], f)
    else:
        z[0] = _at0(z[0], ones)
    if semisolid:
        fb = tt("bound_fraction") * ones
        kb = tt("bound_exchange", 0.0) * ones
        r1bound = 1000.0 / (tt("t1_bound", 1000.0) * ones)
        kab_s, kba_s = kb * fb, kb * (1.0 - fb)
        r1a = 1000.0 / t1
        gs = torch.stack((torch.stack((-kab_s - r1a, kba_s), -1), torch.stack((kab_s, -kba_s - r1bound), -1)), -2)
        split = torch.stack((1.0 - fb, fb), -1)
        z[0] = _at0(z[0], 1.0 - fb)
        zb = _at0(torch.zeros(shape, dtype=_C), fb)
        saturation = torch.as_tensor(events["saturation"], dtype=_R)
        rf_frequency = torch.as_tensor(events["rf_frequency"], dtype=_R)
        knots = torch.as_tensor(lineshape[0], dtype=_R)
        slopes = torch.as_tensor(lineshape[1], dtype=_R)
        step = float(lineshape[2])

        def absorbed(offset):
            # cubic Hermite between the knots around |offset|, clamped to the table
            last = knots.numel() - 1
            scaled = torch.clamp(offset.abs() / step, max=float(last))
            lower = torch.clamp(torch.floor(scaled), max=float(last - 1))
            u = scaled - lower
            i = lower.to(torch.int64)
            return ((2 * u**3 - 3 * u**2 + 1) * knots[i] + (u**3 - 2 * u**2 + u) * step * slopes[i]
                    + (-2 * u**3 + 3 * u**2) * knots[i + 1] + (u**3 - u**2) * step * slopes[i + 1])
    if profile is not None:
        pa, pb, sa_tab, sb_tab = (torch.as_tensor(v, dtype=_C) for v in profile[:4])
        bins = pa.shape[1]
        step_t = float(profile[4]) / (bins - 1)
        where = torch.arange(atoms) % pa.shape[0]

        def pair(theta):
            # cubic Hermite along the flip at each atom's position, clamped
            scaled = torch.clamp(theta / step_t, 0.0, float(bins - 1))
            lower = torch.clamp(torch.floor(scaled), max=float(bins - 2))
            u = scaled - lower
            i = lower.to(torch.int64)
            h = ((2 * u**3 - 3 * u**2 + 1), (u**3 - 2 * u**2 + u) * step_t, (-2 * u**3 + 3 * u**2),
                 (u**3 - u**2) * step_t)

            def read(values, slopes):
                return (h[0] * values[where, i] + h[1] * slopes[where, i] + h[2] * values[where, i + 1]
                        + h[3] * slopes[where, i + 1])

            return read(pa, sa_tab), read(pb, sb_tab)
    samples: list[torch.Tensor | None] = [None] * outputs
    for e in range(len(kind)):
        dt = duration[e]
        damp_t = torch.exp(-(rate * dt)[:, None] * wt[None, :])
        damp_l = torch.exp(-(rate * dt)[:, None] * wl[None, :])
        turn = torch.exp(-2j * torch.pi * b0 * dt)[:, None] * damp_t
        if semisolid:
            e2 = torch.exp(-(1000.0 / t2) * dt)
            fp[0] = fp[0] * turn * e2[:, None]
            fm[0] = fm[0] * turn.conj() * e2[:, None]
            y = torch.linalg.matrix_exp(gs * dt)
            rec = split - (y @ split.unsqueeze(-1)).squeeze(-1)
            l0 = y[:, 0, 0, None] * z[0] + y[:, 0, 1, None] * zb
            l1 = y[:, 1, 0, None] * z[0] + y[:, 1, 1, None] * zb
            z[0] = l0 * damp_l + _at0(l0, rec[:, 0])
            zb = l1 * damp_l + _at0(l1, rec[:, 1])
        elif not exchanging:
            e1 = torch.exp(-(1000.0 / t1) * dt)
            e2 = torch.exp(-(1000.0 / t2) * dt)
            fp[0] = fp[0] * turn * e2[:, None]
            fm[0] = fm[0] * turn.conj() * e2[:, None]
            z[0] = z[0] * (e1[:, None] * damp_l) + _at0(z[0], 1.0 - e1)
        else:
            x = torch.linalg.matrix_exp(gt * dt)
            y = torch.linalg.matrix_exp(gl * dt)
            rec = ((y - torch.eye(2, dtype=_R)) @ settled.unsqueeze(-1)).squeeze(-1)
            a0 = x[:, 0, 0, None] * fp[0] + x[:, 0, 1, None] * fp[1]
            a1 = x[:, 1, 0, None] * fp[0] + x[:, 1, 1, None] * fp[1]
            xc = x.conj()
            c0 = xc[:, 0, 0, None] * fm[0] + xc[:, 0, 1, None] * fm[1]
            c1 = xc[:, 1, 0, None] * fm[0] + xc[:, 1, 1, None] * fm[1]
            l0 = y[:, 0, 0, None] * z[0] + y[:, 0, 1, None] * z[1]
            l1 = y[:, 1, 0, None] * z[0] + y[:, 1, 1, None] * z[1]
            fp = [a0 * turn, a1 * turn]
            fm = [c0 * turn.conj(), c1 * turn.conj()]
            z = [l0 * damp_l + _at0(l0, rec[:, 0]), l1 * damp_l + _at0(l1, rec[:, 1])]
        act = action[e]
        if act & PRE_SHIFT:
            for q in range(pools):
                fp[q], fm[q] = _shift(fp[q], fm[q])
        if kind[e] == 1:
            if act & INVERSION:
                for q in range(pools):
                    z[q] = z[q] * (-ie)[:, None]
            else:
                alpha = (flip[e] * b1)[:, None]
                phi = (phase[e] + b1_phase)[:, None]
                if semisolid and float(saturation[e]) != 0.0:
                    g = absorbed(rf_frequency[e] - b0)
                    zb = zb * torch.exp(saturation[e] * (flip[e] * b1) ** 2 * g)[:, None]
                if profile is not None:
                    a, b = pair(flip[e] * b1)
                    b = b * torch.exp(-1j * (phase[e] + b1_phase))
                    a, b = a[:, None], b[:, None]
                    ca, cb = a.conj(), b.conj()
                    for q in range(pools):
                        x, y, w = fp[q], fm[q], z[q]
                        fp[q] = ca * ca * x - cb * cb * y - 2 * (a * b).conj() * w
                        fm[q] = -b * b * x + a * a * y - 2 * a * b * w
                        z[q] = ca * b * x + a * cb * y + (a.abs() ** 2 - b.abs() ** 2) * w
                    continue_rotation = False
                else:
                    continue_rotation = True
                ca, sa = torch.cos(alpha), torch.sin(alpha)
                p1 = torch.exp(1j * phi)
                p2 = p1 * p1
                t00 = 0.5 * (1 + ca)
                t01 = 0.5 * (1 - ca) * p2
                t02 = -1j * sa * p1
                t10 = t01.conj()
                t12 = 1j * sa * p1.conj()
                t20 = -0.5j * sa * p1.conj()
                t21 = 0.5j * sa * p1
                for q in range(pools if continue_rotation else 0):
                    a, b, c = fp[q], fm[q], z[q]
                    fp[q] = t00 * a + t01 * b + t02 * c
                    fm[q] = t10 * a + t00 * b + t12 * c
                    z[q] = t20 * a + t21 * b + ca * c
        elif kind[e] == 2 and act & RECORD:
            total = sum(fp[q][:, 0] for q in range(pools))
            samples[out_index[e]] = m0 * total * torch.exp(-1j * phase[e])
        if act & POST_SHIFT:
            for q in range(pools):
                fp[q], fm[q] = _shift(fp[q], fm[q])
        if act & SPOIL_AFTER:
            fp = [torch.zeros_like(v) for v in fp]
            fm = [torch.zeros_like(v) for v in fm]
        elif act & SHIFT_AFTER:
            for q in range(pools):
                fp[q], fm[q] = _shift(fp[q], fm[q])
    filled = [s if s is not None else torch.zeros(atoms, dtype=_C) for s in samples]
    if not filled:
        return torch.zeros((atoms, 0), dtype=_C)
    return torch.stack(filled, dim=-1)


def jacobian(
    events: dict[str, Any],
    tissue: dict[str, Any],
    names: tuple[str, ...],
    *,
    states: int,
    outputs: int,
) -> tuple[torch.Tensor, torch.Tensor]:
    """Signal (atoms, outputs) and forward-mode Jacobian (atoms, P, outputs)
    along the per-atom tissue entries ``names`` (FP64, torch.func.jvp)."""
    atoms = max(int(torch.as_tensor(v).numel()) for v in tissue.values() if v is not None)
    base = {k: v for k, v in tissue.items()}
    primals = tuple(torch.as_tensor(base[name], dtype=_R).reshape(-1).expand(atoms).clone() for name in names)

    def run(*values):
        given = dict(base)
        given.update(dict(zip(names, values)))
        return simulate(events, given, states=states, outputs=outputs)

    columns = []
    signal = None
    for index in range(len(names)):
        tangents = tuple(torch.ones_like(v) if j == index else torch.zeros_like(v) for j, v in enumerate(primals))
        signal, column = torch.func.jvp(run, primals, tangents)
        columns.append(column)
    if signal is None:
        signal = run(*primals)
    return signal, torch.stack(columns, dim=1) if columns else torch.zeros((atoms, 0, outputs), dtype=_C)
