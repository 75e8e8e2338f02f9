"""The semisolid (magnetization-transfer) pool on the GPU: the kernel's
signal and Jacobian columns (T1, T2, M0, B1) against the FP64 oracle's
semisolid pool (pinned on the CPU against the reference's own kernel), on an
MRF train with off-resonance and a B1 map and on a refocused train, through
the public simulators and their property names; slice-profiled pulses; and
the derivatives along the sequence through both (the adjoint per event,
the event tangents, loss.backward(), dual tensors and torch.func.jvp)
against FP64 autograd and forward mode of the oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from oracle import packed as P
from tests import _cases

pytestmark = pytest.mark.gpu

TOL = 1e-5    # signal, relative to its largest sample
JTOL = 1e-4   # Jacobian columns, relative to each column's largest entry


def _simulator(name):
    from torchsim_amd.simulators import FSESimulator, MRFSimulator

    if name == "mrf":
        return MRFSimulator(flip=torch.tensor(_cases.benchmark_flips(120), dtype=torch.float32), TR=10.0, states=10)
    # TR far beyond T1: the simulator's recovery factor is one
    return FSESimulator(flip=torch.full((24,), 150.0), ESP=5.0, TR=1e6, states=10)


def _maps(atoms, seed):
    rng = np.random.default_rng(seed)
    t1, t2 = _cases.random_tissue(atoms, seed=seed)
    return {"T1": t1, "T2": t2, "M0": rng.uniform(0.5, 1.0, atoms).astype(np.float32),
            "B1": rng.uniform(0.8, 1.2, atoms).astype(np.float32),
            "B0": rng.uniform(-300.0, 300.0, atoms).astype(np.float32),
            "bound_fraction": rng.uniform(0.05, 0.2, atoms).astype(np.float32),
            "bound_exchange": rng.uniform(5.0, 40.0, atoms).astype(np.float32),
            "bound_T1": rng.uniform(300.0, 1200.0, atoms).astype(np.float32)}


_ORACLE = {"T1": "t1", "T2": "t2", "M0": "m0", "B1": "b1", "B0": "b0", "bound_fraction": "bound_fraction",
           "bound_exchange": "bound_exchange", "bound_T1": "t1_bound"}


@pytest.mark.parametrize("name", ["mrf", "fse"])
def test_signal_and_jacobian_against_the_fp64_oracle(name):
    from torchsim_amd.sequence._lineshape import lineshape_table

    sim = _simulator(name)
    atoms = 48
    maps = _maps(atoms, 17)
    names = ("T1", "T2", "M0", "B1")
    signal, jacobian = sim.jacobian(names, **{k: torch.tensor(v).cuda() for k, v in maps.items()})
    packed = sim.packed(**sim.bound)
    events = oracle.events_of(packed)
    # both trains wind uniformly and sample where nothing is left unrefocused:
    # B0 reaches the samples as a factor of one, and the states (and the
    # lineshape's offset) run without it, as in the reference
    from torchsim_amd.sequence._dephasing import unrefocused_s, winding

    assert winding(packed).analytic and np.abs(unrefocused_s(packed).double().numpy()).max() == 0.0
    maps_states = {k: v for k, v in maps.items() if k != "B0"}
    table = lineshape_table()
    lineshape = (table.values, table.slopes, table.step)
    base = {_ORACLE[k]: torch.tensor(v, dtype=torch.float64) for k, v in maps_states.items()}
    factor = complex(sim.signal_factor) if sim.signal_factor is not None else 1.0

    def run(*values):
        given = dict(base)
        given.update({_ORACLE[k]: v for k, v in zip(names, values)})
        return factor * P.simulate(events, given, states=10, outputs=packed.output_count, lineshape=lineshape)

    primal = tuple(base[_ORACLE[k]] for k in names)
    expect = run(*primal).numpy()
    got = signal.cpu().numpy().reshape(atoms, -1)
    error = np.abs(got - expect).max() / np.abs(expect).max()
    print(f"{name} signal: {error:.2e}")
    assert error < TOL
    for k, n in enumerate(names):
        tangents = tuple(torch.ones(atoms, dtype=torch.float64) if j == k else torch.zeros(atoms, dtype=torch.float64)
                         for j in range(len(names)))
        column = torch.func.jvp(run, primal, tangents)[1].numpy()
        err = np.abs(jacobian[:, k].cpu().numpy() - column).max() / np.abs(column).max()
        print(f"{name} d/d{n}: {err:.2e}")
        assert err < JTOL, n


def test_the_kernel_launch_matches_the_oracle_on_a_refocused_train():
    """The semisolid kernel called directly on the FSE train (the samples
    before the simulator's recovery factor), with columns T1, T2, M0."""
    from torchsim_amd.sequence import _engine, _semisolid
    from torchsim_amd.sequence._lineshape import lineshape_table

    sim = _simulator("fse")
    packed = sim.packed(**sim.bound)
    atoms = 40
    maps = _maps(atoms, 23)
    fields = {"t1_ms": maps["T1"], "t2_ms": maps["T2"], "m0": maps["M0"], "b0_hz": maps["B0"],
              "bound_fraction": maps["bound_fraction"], "bound_exchange_rate": maps["bound_exchange"],
              "t1_bound_ms": maps["bound_T1"]}
    names = ("t1_ms", "t2_ms", "m0")
    tissue = _engine.prepare_tissue({k: torch.tensor(v) for k, v in fields.items()}, torch.device("cuda", 0),
                                    names)
    assert tissue.semisolid
    signal, jacobian = _semisolid.simulate_semisolid(packed, tissue, states=10, columns=names)
    table = lineshape_table()
    lineshape = (table.values, table.slopes, table.step)
    keys = {"T1": "t1", "T2": "t2", "M0": "m0", "B0": "b0", "bound_fraction": "bound_fraction",
            "bound_exchange": "bound_exchange", "bound_T1": "t1_bound"}
    base = {keys[k]: torch.tensor(v, dtype=torch.float64) for k, v in maps.items() if k in keys}
    events = oracle.events_of(packed)

    def run(*values):
        given = dict(base)
        given.update(dict(zip(("t1", "t2", "m0"), values)))
        return P.simulate(events, given, states=10, outputs=packed.output_count, lineshape=lineshape)

    primal = (base["t1"], base["t2"], base["m0"])
    expect = run(*primal).numpy()
    error = np.abs(signal.cpu().numpy() - expect).max() / np.abs(expect).max()
    print(f"fse kernel signal: {error:.2e}")
    assert error < TOL
    for k in range(3):
        tangents = tuple(torch.ones(atoms, dtype=torch.float64) if j == k else torch.zeros(atoms, dtype=torch.float64)
                         for j in range(3))
        column = torch.func.jvp(run, primal, tangents)[1].numpy()
        err = np.abs(jacobian[:, k].cpu().numpy() - column).max() / np.abs(column).max()
        print(f"fse kernel column {names[k]}: {err:.2e}")
        assert err < JTOL


def test_a_zero_bound_fraction_is_the_single_pool_run_bit_for_bit():
    sim = _simulator("mrf")
    maps = {k: torch.tensor(v).cuda() for k, v in _maps(30, 29).items() if not k.startswith("bound")}
    plain = sim.simulate(**maps)
    gated = sim.simulate(**maps, bound_fraction=0.0, bound_exchange=20.0, bound_T1=500.0)
    assert torch.equal(plain, gated)


def _oracle_run(sim, maps, lineshape, profile=None):
    """FP64 signal of the oracle as a function of (T1, T2, M0, B1) per voxel
    (slice-averaged where a transition table is given; B0 left out, as the
    uniform winding leaves it a factor of one at the samples)."""
    packed = sim.packed(**sim.bound)
    events = oracle.events_of(packed)
    names = ("T1", "T2", "M0", "B1")
    voxels = len(maps["T1"])
    positions = 1 if profile is None else profile[0].shape[0]
    base = {_ORACLE[k]: torch.tensor(np.repeat(v, positions), dtype=torch.float64) for k, v in maps.items()
            if k != "B0"}
    factor = complex(sim.signal_factor) if sim.signal_factor is not None else 1.0

    def run(*values):
        given = dict(base)
        given.update({_ORACLE[k]: v.repeat_interleave(positions) for k, v in zip(names, values)})
        rows = P.simulate(events, given, states=10, outputs=packed.output_count, lineshape=lineshape,
                          profile=profile)
        return factor * rows.reshape(voxels, positions, -1).mean(1)

    return run, tuple(torch.tensor(maps[k], dtype=torch.float64) for k in names)


@pytest.mark.parametrize("name", ["mrf", "fse_profiled"])
def test_gradients_reach_the_tissue_maps(name):
    """loss.backward() through the semisolid kernel (and a profiled pulse on
    top of the pool) reaches T1, T2, M0 and B1: against FP64 autograd of the
    oracle, each map's gradient within 1e-4 of its largest entry."""
    from torchsim_amd.sequence._lineshape import lineshape_table
    from torchsim_amd.sequence._transition import exact_slice_profile, sinc_pulse, transition_table

    sim = _simulator(name.split("_")[0])
    profile = None
    if name.endswith("profiled"):
        pulse, where = sinc_pulse(2e-3, 4.0, raster_s=1e-5), exact_slice_profile(5, bins=32)
        sim = sim.bind(pulse=pulse, across_slice=where)
        table = transition_table(pulse, where.positions(), bins=where.bins, theta_max=where.theta_max)
        profile = (table.a.numpy(), table.b.numpy(), table.slope_a.numpy(), table.slope_b.numpy(), table.theta_max)
    atoms = 24
    maps = _maps(atoms, 43)
    ls = lineshape_table()
    run, primal = _oracle_run(sim, maps, (ls.values, ls.slopes, ls.step), profile)
    names = ("T1", "T2", "M0", "B1")
    leaves = {k: torch.tensor(maps[k]).cuda().requires_grad_(True) for k in names}
    rest = {k: torch.tensor(v).cuda() for k, v in maps.items() if k not in names}
    signal = sim.simulate(**leaves, **rest)
    rng = np.random.default_rng(5)
    target = rng.standard_normal(tuple(signal.shape)) + 1j * rng.standard_normal(tuple(signal.shape))
    loss = (signal - torch.tensor(target, dtype=torch.complex64).cuda()).abs().pow(2).sum()
    loss.backward()
    leaves64 = tuple(p.clone().requires_grad_(True) for p in primal)
    expect = run(*leaves64)
    (expect - torch.tensor(target)).abs().pow(2).sum().backward()
    for k, leaf in zip(names, leaves64):
        got = leaves[k].grad.cpu().double().numpy()
        want = leaf.grad.numpy()
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"{name} dloss/d{k}: {err:.2e}")
        assert err < JTOL, k


def test_forward_mode_along_a_mixed_tissue_direction():
    """torch.func.jvp and dual tensors along (T1, B1) together through the
    semisolid kernel, against the FP64 oracle's jvp."""
    import torch.autograd.forward_ad as fwAD
    from torchsim_amd.sequence._lineshape import lineshape_table

    sim = _simulator("mrf")
    atoms = 20
    maps = _maps(atoms, 47)
    ls = lineshape_table()
    run, primal = _oracle_run(sim, maps, (ls.values, ls.slopes, ls.step))
    rng = np.random.default_rng(9)
    d_t1 = rng.uniform(-50.0, 50.0, atoms).astype(np.float32)
    d_b1 = rng.uniform(-0.1, 0.1, atoms).astype(np.float32)
    zero = torch.zeros(atoms, dtype=torch.float64)
    want = torch.func.jvp(run, primal, (torch.tensor(d_t1, dtype=torch.float64), zero, zero,
                                        torch.tensor(d_b1, dtype=torch.float64)))[1].numpy()
    rest = {k: torch.tensor(v).cuda() for k, v in maps.items() if k not in ("T1", "B1")}

    def simulate(t1, b1):
        return sim.simulate(T1=t1, B1=b1, **rest)

    t1, b1 = torch.tensor(maps["T1"]).cuda(), torch.tensor(maps["B1"]).cuda()
    _, got = torch.func.jvp(simulate, (t1, b1), (torch.tensor(d_t1).cuda(), torch.tensor(d_b1).cuda()))
    err = np.abs(got.cpu().numpy() - want).max() / np.abs(want).max()
    print(f"torch.func.jvp: {err:.2e}")
    assert err < JTOL
    with fwAD.dual_level():
        dual = simulate(fwAD.make_dual(t1, torch.tensor(d_t1).cuda()), fwAD.make_dual(b1, torch.tensor(d_b1).cuda()))
        tangent = fwAD.unpack_dual(dual).tangent
    err = np.abs(tangent.cpu().numpy() - want).max() / np.abs(want).max()
    print(f"dual tensors: {err:.2e}")
    assert err < JTOL


def test_derivatives_the_pool_does_not_carry_are_refused():
    """A Jacobian's gradient to the flips, a Jacobian along the flips' dual
    tangents, and the bound pool's own properties carry no derivative
    through the semisolid kernel: asking raises."""
    import torch.autograd.forward_ad as fwAD
    from torchsim_amd.simulators import MRFSimulator

    maps = {k: torch.tensor(v).cuda() for k, v in _maps(8, 31).items()}
    flips = torch.tensor(_cases.benchmark_flips(40), dtype=torch.float32, requires_grad=True)
    with pytest.raises(NotImplementedError, match="not differentiated again"):
        MRFSimulator(flip=flips, TR=10.0, states=10).jacobian(("T1",), **maps)
    with fwAD.dual_level():
        dual = fwAD.make_dual(flips.detach(), torch.ones_like(flips))
        with pytest.raises(NotImplementedError, match="not differentiated again"):
            MRFSimulator(flip=dual, TR=10.0, states=10).jacobian(("T1",), **maps)
    sim = _simulator("mrf")
    bound = dict(maps)
    bound["bound_fraction"] = bound["bound_fraction"].clone().requires_grad_(True)
    with pytest.raises(NotImplementedError, match="bound_fraction"):
        sim.simulate(**bound)


# --- slice-profiled pulses ---


def _profiled_expected(sim, maps, names, table, lineshape=None):
    """FP64: the oracle at every tabulated position (one row per voxel and
    position), averaged over positions, with its forward-mode columns."""
    packed = sim.packed(**sim.bound)
    events = oracle.events_of(packed)
    positions = table.points
    profile = (table.a.numpy(), table.b.numpy(), table.slope_a.numpy(), table.slope_b.numpy(), table.theta_max)
    base = {_ORACLE[k]: torch.tensor(np.repeat(v, positions), dtype=torch.float64) for k, v in maps.items()}
    factor = complex(sim.signal_factor) if sim.signal_factor is not None else 1.0
    voxels = len(maps["T1"])

    def run(*values):
        given = dict(base)
        given.update({_ORACLE[k]: v.repeat_interleave(positions) for k, v in zip(names, values)})
        rows = P.simulate(events, given, states=10, outputs=packed.output_count, lineshape=lineshape,
                          profile=profile)
        return factor * rows.reshape(voxels, positions, -1).mean(1)

    primal = tuple(torch.tensor(maps[k], dtype=torch.float64) for k in names)
    signal = run(*primal).numpy()
    columns = []
    for k in range(len(names)):
        tangents = tuple(torch.ones(voxels, dtype=torch.float64) if j == k else torch.zeros(voxels, dtype=torch.float64)
                         for j in range(len(names)))
        columns.append(torch.func.jvp(run, primal, tangents)[1].numpy())
    return signal, columns


@pytest.mark.parametrize("name", ["mrf", "fse", "fse_semisolid"])
def test_profiled_pulses_against_the_fp64_oracle(name):
    """A windowed-sinc pulse worked out at 7 slice positions (the table the
    oracle reads too): signal and T1/T2/M0/B1 columns of the slice average,
    with a B1 map (and, for one case, a semisolid pool, whose saturation is
    the pulse's own deposited power)."""
    from torchsim_amd.sequence._lineshape import lineshape_table
    from torchsim_amd.sequence._transition import exact_slice_profile, sinc_pulse, transition_table

    pulse = sinc_pulse(2e-3, 4.0, raster_s=1e-5)
    profile = exact_slice_profile(7, bins=32)
    sim = _simulator(name.split("_")[0]).bind(pulse=pulse, across_slice=profile)
    atoms = 20
    maps = {k: v for k, v in _maps(atoms, 37).items() if k in ("T1", "T2", "M0", "B1")}
    lineshape = None
    if name == "fse_semisolid":
        extra = _maps(atoms, 38)
        maps.update({k: extra[k] for k in ("bound_fraction", "bound_exchange", "bound_T1")})
        table_ls = lineshape_table()
        lineshape = (table_ls.values, table_ls.slopes, table_ls.step)
    names = ("T1", "T2", "M0", "B1")
    signal, jacobian = sim.jacobian(names, **{k: torch.tensor(v).cuda() for k, v in maps.items()})
    table = transition_table(pulse, profile.positions(), bins=profile.bins, theta_max=profile.theta_max)
    expect, columns = _profiled_expected(sim, maps, names, table, lineshape)
    error = np.abs(signal.cpu().numpy() - expect).max() / np.abs(expect).max()
    print(f"{name} profiled signal: {error:.2e}")
    assert error < TOL
    for k, n in enumerate(names):
        err = np.abs(jacobian[:, k].cpu().numpy() - columns[k]).max() / np.abs(columns[k]).max()
        print(f"{name} profiled d/d{n}: {err:.2e}")
        assert err < JTOL, n


def test_a_tabulated_hard_pulse_is_the_instant_rotation():
    """A rectangle without slice selection asked for at the slice centre
    runs the profiled kernel through its table, and lands on the hard pulse's
    answer (to the table's interpolation)."""
    from torchsim_amd.sequence._transition import RfPulse

    sim = _simulator("mrf")
    maps = {k: torch.tensor(v).cuda() for k, v in _maps(16, 41).items() if k in ("T1", "T2", "M0")}
    instant = sim.simulate(**maps)
    tabulated = sim.simulate(**maps, pulse=RfPulse(np.ones(20), 1e-5, 0.0), across_slice=1)
    error = float((tabulated - instant).abs().max() / instant.abs().max())
    print(f"hard pulse through its table: {error:.2e}")
    assert error < 2e-5


# --- reverse mode to the sequence through the pool ---

SEQ_TOL = 2e-5  # per gradient, relative to its largest entry


def _pool_tissue(atoms, seed):
    rng = np.random.default_rng(seed)
    t1, t2 = _cases.random_tissue(atoms, seed=seed)
    return {"t1": t1, "t2": t2, "m0": rng.uniform(0.5, 1.0, atoms), "b1": rng.uniform(0.8, 1.2, atoms),
            "b0": rng.uniform(-200.0, 200.0, atoms), "bound_fraction": rng.uniform(0.05, 0.2, atoms),
            "bound_exchange": rng.uniform(5.0, 40.0, atoms), "t1_bound": rng.uniform(300.0, 1200.0, atoms)}


def test_the_adjoint_kernel_against_fp64_autograd_per_event():
    """tsim_semisolid_vjp on a phase-cycled MRF train with B0, a B1 map, the
    pulses played 800 Hz off resonance and their saturation scaled 300-fold
    (so that the bound pool's term moves the flip gradient by percents):
    every event's flip, phase and duration gradient of Re(sum conj(W) s)
    against FP64 autograd of the oracle on the same events."""
    from torchsim_amd.sequence import _engine, _semisolid
    from torchsim_amd.sequence._lineshape import lineshape_table
    from torchsim_amd.simulators import MRFSimulator

    atoms, pulses, states = 20, 80, 10
    rng = np.random.default_rng(5)
    flips, trs, phases = rng.uniform(5.0, 60.0, pulses), rng.uniform(8.0, 15.0, pulses), rng.uniform(0, 360, pulses)
    packed = MRFSimulator(states=states).packed(flip=torch.tensor(flips, dtype=torch.float32),
                                                 TR=torch.tensor(trs, dtype=torch.float32),
                                                 phases=torch.tensor(phases, dtype=torch.float32))
    import dataclasses as dc

    packed = dc.replace(packed, saturation=packed.saturation * 300.0,
                        rf_frequency_hz=torch.where(packed.kind == 1, 800.0, 0.0).to(torch.float32))
    tissue = _pool_tissue(atoms, 6)
    fields = {"t1_ms": tissue["t1"], "t2_ms": tissue["t2"], "m0": tissue["m0"], "b1": tissue["b1"],
              "b0_hz": tissue["b0"], "bound_fraction": tissue["bound_fraction"],
              "bound_exchange_rate": tissue["bound_exchange"], "t1_bound_ms": tissue["t1_bound"]}
    prepared = _engine.prepare_tissue({k: torch.tensor(v, dtype=torch.float32) for k, v in fields.items()},
                                      torch.device("cuda", 0), ())
    assert prepared.semisolid
    weight = rng.normal(size=(atoms, packed.output_count)) + 1j * rng.normal(size=(atoms, packed.output_count))
    signal, _ = _semisolid.simulate_semisolid(packed, prepared, states=states, factor=1j)
    grads = _semisolid.adjoint_semisolid(packed, prepared, states=states, factor=1j,
                                         cotangent=torch.tensor(weight, dtype=torch.complex64).cuda())
    table = lineshape_table()
    events = oracle.events_of(packed)
    leaves = {k: torch.tensor(events[k], dtype=torch.float64, requires_grad=True) for k in ("flip", "phase", "duration")}
    given = dict(events)
    given.update(leaves)
    ref = 1j * P.simulate(given, {k: torch.tensor(v, dtype=torch.float64) for k, v in tissue.items()}, states=states,
                          outputs=packed.output_count, lineshape=(table.values, table.slopes, table.step))
    error = np.abs(signal.cpu().numpy() - ref.detach().numpy()).max() / np.abs(ref.detach().numpy()).max()
    print(f"signal: {error:.2e}")
    assert error < 1e-5
    (ref * torch.tensor(weight).conj()).real.sum().backward()
    for name, got in zip(("flip", "phase", "duration"), grads):
        want = leaves[name].grad.numpy()
        err = np.abs(got.cpu().numpy() - want).max() / np.abs(want).max()
        print(f"d/d{name}: {err:.2e}")
        assert err < SEQ_TOL, name
    # the bound pool's saturation term matters here: without it the flip gradient misses
    plain = dict(given)
    plain["saturation"] = np.zeros_like(events["saturation"])
    for leaf in leaves.values():
        leaf.grad = None
    ref2 = 1j * P.simulate(plain, {k: torch.tensor(v, dtype=torch.float64) for k, v in tissue.items()},
                           states=states, outputs=packed.output_count,
                           lineshape=(table.values, table.slopes, table.step))
    (ref2 * torch.tensor(weight).conj()).real.sum().backward()
    miss = np.abs(grads[0].cpu().numpy() - leaves["flip"].grad.numpy()).max() / np.abs(leaves["flip"].grad.numpy()).max()
    print(f"flip gradient against a saturation-free oracle: {miss:.2e}")
    assert miss > 1e-3


@pytest.mark.parametrize("name", ["mrf", "fse"])
def test_loss_backward_reaches_the_protocol_through_the_pool(name):
    """loss.backward() through a simulator with a semisolid pool reaches its
    flips and TRs (MRF: and phases) and the T1/T2 maps, against FP64 autograd
    of the oracle through the same differentiable packing."""
    from torchsim_amd.sequence._lineshape import lineshape_table
    from torchsim_amd.simulators import FSESimulator, MRFSimulator

    atoms = 16
    rng = np.random.default_rng(11)
    tissue = _pool_tissue(atoms, 12)
    tissue.pop("b0")  # B0 in the states only where the winding is not uniform; these trains wind uniformly
    if name == "mrf":
        pulses = 60
        values = {"flip": rng.uniform(5.0, 60.0, pulses), "TR": rng.uniform(8.0, 15.0, pulses),
                  "phases": rng.uniform(0.0, 360.0, pulses)}

        def build(**kw):
            return MRFSimulator(states=10, **kw)
    else:
        values = {"flip": rng.uniform(120.0, 180.0, 16)}

        def build(**kw):
            return FSESimulator(ESP=5.0, TR=1e6, states=10, **kw)
    ours = {k: torch.tensor(v, dtype=torch.float32, requires_grad=True) for k, v in values.items()}
    maps = {k: torch.tensor(tissue[k], dtype=torch.float32, device="cuda", requires_grad=(k in ("t1", "t2")))
            for k in tissue}
    public = {"t1": "T1", "t2": "T2", "m0": "M0", "b1": "B1", "bound_fraction": "bound_fraction",
              "bound_exchange": "bound_exchange", "t1_bound": "bound_T1"}
    sim = build(**ours)
    sig = sim.simulate(**{public[k]: v for k, v in maps.items()})
    weight = rng.normal(size=tuple(sig.shape)) + 1j * rng.normal(size=tuple(sig.shape))
    (sig * torch.tensor(weight, dtype=torch.complex64).cuda().conj()).real.sum().backward()

    table = lineshape_table()
    theirs = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in values.items()}
    sim64 = build(**theirs)
    packed = sim64.packed(**sim64.bound)
    events = oracle.events_of(packed)
    events.update(duration=packed.duration.double(), flip=packed.flip.double(), phase=packed.phase.double())
    maps64 = {k: torch.tensor(v, dtype=torch.float64, requires_grad=(k in ("t1", "t2"))) for k, v in tissue.items()}
    factor = complex(sim64.signal_factor) if sim64.signal_factor is not None else 1.0
    ref = factor * P.simulate(events, maps64, states=10, outputs=packed.output_count,
                              lineshape=(table.values, table.slopes, table.step))
    (ref * torch.tensor(weight).conj()).real.sum().backward()
    for key in values:
        err = np.abs(ours[key].grad.numpy() - theirs[key].grad.numpy()).max() / np.abs(theirs[key].grad.numpy()).max()
        print(f"{name} d/d{key}: {err:.2e}")
        assert err < SEQ_TOL, key
    for key in ("t1", "t2"):
        err = np.abs(maps[key].grad.cpu().numpy() - maps64[key].grad.numpy()).max() / np.abs(maps64[key].grad.numpy()).max()
        print(f"{name} d/d{key}: {err:.2e}")
        assert err < JTOL, key


@pytest.mark.parametrize("bound", [False, True], ids=["free", "semisolid"])
def test_the_profiled_adjoint_against_fp64_autograd_per_event(bound):
    """tsim_profiled_vjp: a windowed sinc worked out at 5 slice positions on
    a phase-cycled MRF train with a B1 map (and a semisolid pool at 300x
    saturation): every event's flip, phase and duration gradient of the
    slice-averaged loss against FP64 autograd of the oracle on the same
    table."""
    import dataclasses as dc

    from torchsim_amd.sequence import _engine, _semisolid
    from torchsim_amd.sequence._lineshape import lineshape_table
    from torchsim_amd.sequence._transition import sinc_pulse, transition_table
    from torchsim_amd.simulators import MRFSimulator

    voxels, pulses, states, positions = 8, 50, 10, 5
    rng = np.random.default_rng(21)
    packed = MRFSimulator(states=states).packed(flip=torch.tensor(rng.uniform(5.0, 60.0, pulses), dtype=torch.float32),
                                                 TR=torch.tensor(rng.uniform(8.0, 15.0, pulses), dtype=torch.float32),
                                                 phases=torch.tensor(rng.uniform(0, 360, pulses), dtype=torch.float32))
    if bound:
        packed = dc.replace(packed, saturation=packed.saturation * 300.0)
    table = transition_table(sinc_pulse(2e-3, 4.0, raster_s=1e-5), np.linspace(-1.0, 1.0, positions), bins=32)
    tissue = _pool_tissue(voxels, 22)
    tissue.pop("b0")
    if not bound:
        for key in ("bound_fraction", "bound_exchange", "t1_bound"):
            tissue.pop(key)
    names = {"t1": "t1_ms", "t2": "t2_ms", "m0": "m0", "b1": "b1", "bound_fraction": "bound_fraction",
             "bound_exchange": "bound_exchange_rate", "t1_bound": "t1_bound_ms"}
    prepared = _engine.prepare_tissue({names[k]: torch.tensor(v, dtype=torch.float32) for k, v in tissue.items()},
                                      torch.device("cuda", 0), ())
    weight = rng.normal(size=(voxels, packed.output_count)) + 1j * rng.normal(size=(voxels, packed.output_count))
    signal, _ = _semisolid.simulate_profiled(packed, prepared, table, states=states, factor=1j)
    grads = _semisolid.adjoint_semisolid(packed, prepared, states=states, factor=1j, table=table,
                                         cotangent=torch.tensor(weight, dtype=torch.complex64).cuda())
    events = oracle.events_of(packed)
    leaves = {k: torch.tensor(events[k], dtype=torch.float64, requires_grad=True) for k in ("flip", "phase", "duration")}
    given = dict(events)
    given.update(leaves)
    rows = {k: torch.tensor(np.repeat(v, positions), dtype=torch.float64) for k, v in tissue.items()}
    ls = lineshape_table()
    profile = (table.a.numpy(), table.b.numpy(), table.slope_a.numpy(), table.slope_b.numpy(), table.theta_max)
    out = 1j * P.simulate(given, rows, states=states, outputs=packed.output_count, profile=profile,
                          lineshape=(ls.values, ls.slopes, ls.step) if bound else None)
    ref = out.reshape(voxels, positions, -1).mean(1)
    error = np.abs(signal.cpu().numpy() - ref.detach().numpy()).max() / np.abs(ref.detach().numpy()).max()
    print(f"profiled signal: {error:.2e}")
    assert error < 1e-5
    (ref * torch.tensor(weight).conj()).real.sum().backward()
    for name, got in zip(("flip", "phase", "duration"), grads):
        want = leaves[name].grad.numpy()
        err = np.abs(got.cpu().numpy() - want).max() / np.abs(want).max()
        print(f"profiled d/d{name}: {err:.2e}")
        assert err < SEQ_TOL, name


def test_loss_backward_reaches_a_profiled_protocol():
    """loss.backward() through MRFSimulator(pulse=, across_slice=) reaches the
    flips and TRs, against FP64 autograd of the oracle on the same table."""
    from torchsim_amd.sequence._transition import exact_slice_profile, sinc_pulse, transition_table
    from torchsim_amd.simulators import MRFSimulator

    rng = np.random.default_rng(31)
    pulses, voxels = 40, 6
    values = {"flip": rng.uniform(5.0, 60.0, pulses), "TR": rng.uniform(8.0, 15.0, pulses)}
    pulse, where = sinc_pulse(2e-3, 4.0, raster_s=1e-5), exact_slice_profile(5, bins=32)
    t1, t2 = _cases.random_tissue(voxels, seed=32)
    ours = {k: torch.tensor(v, dtype=torch.float32, requires_grad=True) for k, v in values.items()}
    sim = MRFSimulator(states=10, pulse=pulse, across_slice=where, **ours)
    sig = sim.simulate(T1=torch.tensor(t1).cuda(), T2=torch.tensor(t2).cuda())
    weight = rng.normal(size=tuple(sig.shape)) + 1j * rng.normal(size=tuple(sig.shape))
    (sig * torch.tensor(weight, dtype=torch.complex64).cuda().conj()).real.sum().backward()

    table = transition_table(pulse, where.positions(), bins=where.bins, theta_max=where.theta_max)
    theirs = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in values.items()}
    sim64 = MRFSimulator(states=10, pulse=pulse, across_slice=where, **theirs)
    packed = sim64.packed(**sim64.bound)
    events = oracle.events_of(packed)
    events.update(duration=packed.duration.double(), flip=packed.flip.double(), phase=packed.phase.double())
    positions = table.points
    rows = {"t1": torch.tensor(np.repeat(t1, positions), dtype=torch.float64),
            "t2": torch.tensor(np.repeat(t2, positions), dtype=torch.float64)}
    profile = (table.a.numpy(), table.b.numpy(), table.slope_a.numpy(), table.slope_b.numpy(), table.theta_max)
    out = complex(sim64.signal_factor) * P.simulate(events, rows, states=10, outputs=packed.output_count,
                                                   profile=profile)
    ref = out.reshape(voxels, positions, -1).mean(1)
    (ref * torch.tensor(weight).conj()).real.sum().backward()
    for key in values:
        err = np.abs(ours[key].grad.numpy() - theirs[key].grad.numpy()).max() / np.abs(theirs[key].grad.numpy()).max()
        print(f"profiled protocol d/d{key}: {err:.2e}")
        assert err < SEQ_TOL, key


# --- forward mode along the events ---


def _oracle_event_jvp(events, rows, states, outputs, directions, **kw):
    """FP64 forward mode of the oracle along per-event flip, phase and
    duration directions: (signal, tangent)."""
    def run(flip, phase, duration):
        given = dict(events)
        given.update(flip=flip, phase=phase, duration=duration)
        return P.simulate(given, rows, states=states, outputs=outputs, **kw)

    primal = tuple(torch.as_tensor(np.asarray(events[k]), dtype=torch.float64) for k in ("flip", "phase", "duration"))
    tangent = tuple(torch.as_tensor(d, dtype=torch.float64) for d in directions)
    return torch.func.jvp(run, primal, tangent)


@pytest.mark.parametrize("case", ["semisolid", "profiled_free", "profiled_semisolid"])
def test_the_event_tangent_kernel_against_fp64_forward_mode(case):
    """tsim_semisolid_event_jvp on a phase-cycled MRF train with a B1 map
    (semisolid: B0, pulses 800 Hz off resonance, saturation scaled 300-fold;
    profiled: a windowed sinc at 5 slice positions): the signal's tangent
    along each of the flips, the phases and the durations alone, and along
    all three together, against FP64 forward mode of the oracle."""
    import dataclasses as dc

    from torchsim_amd.sequence import _engine, _semisolid
    from torchsim_amd.sequence._lineshape import lineshape_table
    from torchsim_amd.sequence._transition import sinc_pulse, transition_table

    from torchsim_amd.simulators import MRFSimulator

    voxels, pulses, states = 12, 60, 10
    rng = np.random.default_rng(41)
    packed = MRFSimulator(states=states).packed(flip=torch.tensor(rng.uniform(5.0, 60.0, pulses), dtype=torch.float32),
                                                 TR=torch.tensor(rng.uniform(8.0, 15.0, pulses), dtype=torch.float32),
                                                 phases=torch.tensor(rng.uniform(0, 360, pulses), dtype=torch.float32))
    bound = case != "profiled_free"
    profiled = case.startswith("profiled")
    if bound:
        packed = dc.replace(packed, saturation=packed.saturation * 300.0)
    if case == "semisolid":
        packed = dc.replace(packed, rf_frequency_hz=torch.where(packed.kind == 1, 800.0, 0.0).to(torch.float32))
    tissue = _pool_tissue(voxels, 42)
    if profiled:
        tissue.pop("b0")
    if not bound:
        for key in ("bound_fraction", "bound_exchange", "t1_bound"):
            tissue.pop(key)
    names = {"t1": "t1_ms", "t2": "t2_ms", "m0": "m0", "b1": "b1", "b0": "b0_hz", "bound_fraction": "bound_fraction",
             "bound_exchange": "bound_exchange_rate", "t1_bound": "t1_bound_ms"}
    prepared = _engine.prepare_tissue({names[k]: torch.tensor(v, dtype=torch.float32) for k, v in tissue.items()},
                                      torch.device("cuda", 0), ())
    table = transition_table(sinc_pulse(2e-3, 4.0, raster_s=1e-5), np.linspace(-1.0, 1.0, 5), bins=32) if profiled else None
    positions = 1 if table is None else table.points
    events = oracle.events_of(packed)
    rows = {k: torch.tensor(np.repeat(v, positions), dtype=torch.float64) for k, v in tissue.items()}
    ls = lineshape_table()
    kw = {"lineshape": (ls.values, ls.slopes, ls.step) if bound else None}
    if table is not None:
        kw["profile"] = (table.a.numpy(), table.b.numpy(), table.slope_a.numpy(), table.slope_b.numpy(),
                         table.theta_max)
    count = len(events["flip"])
    each = [rng.normal(size=count), rng.normal(size=count), rng.uniform(-0.5, 0.5, count)]
    zero = np.zeros(count)
    trials = {"flip": (each[0], zero, zero), "phase": (zero, each[1], zero), "duration": (zero, zero, each[2]),
              "all": tuple(each)}
    for label, directions in trials.items():
        _, want = _oracle_event_jvp(events, rows, states, packed.output_count, directions, **kw)
        want = (1j * want).reshape(voxels, positions, -1).mean(1).numpy()
        got = _semisolid.event_tangent(packed, prepared, table, states=states, factor=1j,
                                       **{f"{k}_tangent": torch.tensor(d, dtype=torch.float32)
                                          for k, d in zip(("flip", "phase", "duration"), directions)})
        err = np.abs(got.cpu().numpy() - want).max() / np.abs(want).max()
        print(f"{case} along {label}: {err:.2e}")
        assert err < SEQ_TOL, label


@pytest.mark.parametrize("name", ["mrf", "profiled"])
def test_forward_mode_along_the_protocol_through_the_pool(name):
    """torch.func.jvp and dual tensors along the flips, TRs and phases of an
    MRF protocol with a semisolid pool (and, profiled, a windowed sinc across
    the slice), with a T1 direction beside them, against FP64 forward mode of
    the oracle through the same differentiable packing."""
    import torch.autograd.forward_ad as fwAD
    from torchsim_amd.sequence._lineshape import lineshape_table
    from torchsim_amd.sequence._transition import exact_slice_profile, sinc_pulse, transition_table
    from torchsim_amd.simulators import MRFSimulator

    atoms, pulses = 10, 40
    rng = np.random.default_rng(51)
    tissue = _pool_tissue(atoms, 52)
    tissue.pop("b0")
    extra = {}
    if name == "profiled":
        extra = {"pulse": sinc_pulse(2e-3, 4.0, raster_s=1e-5), "across_slice": exact_slice_profile(5, bins=32)}
    values = {"flip": rng.uniform(5.0, 60.0, pulses), "TR": rng.uniform(8.0, 15.0, pulses),
              "phases": rng.uniform(0.0, 360.0, pulses)}
    directions = {"flip": rng.normal(size=pulses), "TR": rng.normal(size=pulses),
                  "phases": rng.normal(size=pulses) * 5.0}
    d_t1 = rng.uniform(-50.0, 50.0, atoms)
    public = {"t1": "T1", "t2": "T2", "m0": "M0", "b1": "B1", "bound_fraction": "bound_fraction",
              "bound_exchange": "bound_exchange", "t1_bound": "bound_T1"}
    rest = {public[k]: torch.tensor(v, dtype=torch.float32, device="cuda") for k, v in tissue.items() if k != "t1"}

    def simulate(flip, tr, phases, t1):
        return MRFSimulator(states=10, flip=flip, TR=tr, phases=phases, **extra).simulate(T1=t1, **rest)

    primal = tuple(torch.tensor(values[k], dtype=torch.float32) for k in values) + (
        torch.tensor(tissue["t1"], dtype=torch.float32, device="cuda"),)
    tangent = tuple(torch.tensor(directions[k], dtype=torch.float32) for k in values) + (
        torch.tensor(d_t1, dtype=torch.float32, device="cuda"),)

    ls = lineshape_table()
    kw = {"lineshape": (ls.values, ls.slopes, ls.step)}
    positions = 1
    if name == "profiled":
        where = extra["across_slice"]
        table = transition_table(extra["pulse"], where.positions(), bins=where.bins, theta_max=where.theta_max)
        positions = table.points
        kw["profile"] = (table.a.numpy(), table.b.numpy(), table.slope_a.numpy(), table.slope_b.numpy(),
                         table.theta_max)

    def reference(flip, tr, phases, t1):
        sim64 = MRFSimulator(states=10, flip=flip, TR=tr, phases=phases, **extra)
        packed = sim64.packed(**sim64.bound)
        events = oracle.events_of(packed)
        events.update(duration=packed.duration.double(), flip=packed.flip.double(), phase=packed.phase.double())
        rows = {k: torch.tensor(np.repeat(v, positions), dtype=torch.float64) for k, v in tissue.items() if k != "t1"}
        rows["t1"] = t1.repeat_interleave(positions)
        out = complex(sim64.signal_factor) * P.simulate(events, rows, states=10, outputs=packed.output_count, **kw)
        return out.reshape(atoms, positions, -1).mean(1)

    with fwAD.dual_level():
        want = reference(*(fwAD.make_dual(torch.tensor(values[k], dtype=torch.float64),
                                          torch.tensor(directions[k], dtype=torch.float64)) for k in values),
                         fwAD.make_dual(torch.tensor(tissue["t1"], dtype=torch.float64),
                                        torch.tensor(d_t1, dtype=torch.float64)))
        want = fwAD.unpack_dual(want).tangent.numpy()
    _, got = torch.func.jvp(simulate, primal, tangent)
    err = np.abs(got.cpu().numpy() - want).max() / np.abs(want).max()
    print(f"{name} torch.func.jvp: {err:.2e}")
    assert err < SEQ_TOL
    with fwAD.dual_level():
        dual = simulate(*(fwAD.make_dual(p, t) for p, t in zip(primal, tangent)))
        tangent_out = fwAD.unpack_dual(dual).tangent
    err = np.abs(tangent_out.cpu().numpy() - want).max() / np.abs(want).max()
    print(f"{name} dual tensors: {err:.2e}")
    assert err < SEQ_TOL
