"""Local time stepping: the stepper's levels, rates, lists and passes against the
reference's own unit tests (src/time/multirate.rs:1349-1805), and whole
multirate runs of the restatement against the reference's gates
(tests/local_time_stepping_test.rs): one level is the base scheme bit for bit,
several levels conserve mass to rounding and save RHS work, a lake at rest with
an island stays at rest across levels."""
import math

import numpy as np
import pytest

from dgcom_amd import cases
from dgcom_amd import multirate as M
import oracle as O


def row_neighbours(n):
    """Face neighbours of a row of n elements (south, east, north, west)."""
    mesh = cases.structured_mesh(0.0, float(n), 0.0, 1.0, n, 1)
    return M.neighbours(mesh.links)


def test_levels_are_powers_of_two_below_the_coarse_step():
    edt = [1.0, 0.3, 0.26, 0.1, math.inf, 0.55]
    dt, finest, levels = M.assign_levels(edt, 8)
    assert finest == 3 and dt == 0.8
    assert list(levels) == [0, 2, 2, 3, 0, 1]
    for l, dk in zip(levels, edt):
        dt_l = dt / (1 << int(l))
        assert dt_l <= dk and (l == 0 or 2.0 * dt_l > dk)


def test_levels_respect_the_caps():
    dt, finest, levels = M.assign_levels([1.0, 0.1], 2)
    assert (dt, finest, list(levels)) == (0.4, 2, [0, 2])
    dt, finest, levels = M.assign_levels([1.0, 0.1], 8, 0.25)
    assert (dt, finest, list(levels)) == (0.2, 1, [0, 1])
    dt, finest, levels = M.assign_levels([1.0, 0.1], 0)
    assert (dt, finest, list(levels)) == (0.1, 0, [0, 0])
    dt, finest, levels = M.assign_levels([math.inf] * 3, 4)
    assert (dt, finest, list(levels)) == (math.inf, 0, [0, 0, 0])


def test_lists_hold_the_elements_at_or_above_each_key_in_order():
    lists = M.at_or_above(np.array([0, 2, 1, 2, 0, 3], dtype=np.uint8), 3)
    assert [list(l) for l in lists] == [[0, 1, 2, 3, 4, 5], [1, 2, 3, 5], [1, 3, 5], [5]]
    lists = M.at_or_above(np.array([0, 2], dtype=np.uint8), 1)
    assert [list(l) for l in lists] == [[0, 1], [1]]


def test_depth_is_the_coarsest_aligned_level():
    assert [M.depth(p, 3) for p in range(9)] == [0, 3, 2, 3, 1, 3, 2, 3, 0]
    assert M.depth(0, 0) == 0 and M.depth(1, 0) == 0


def test_rates_spread_one_face_per_stage():
    nb = row_neighbours(6)
    r1, _ = M.spread_to_neighbours(nb, np.array([0, 0, 0, 0, 0, 2], dtype=np.uint8), 0)
    r2, _ = M.spread_to_neighbours(nb, r1, 0)
    r3, _ = M.spread_to_neighbours(nb, r2, 0)
    assert list(r1) == [0, 0, 0, 0, 2, 2]
    assert list(r2) == [0, 0, 0, 2, 2, 2]
    assert list(r3) == [0, 0, 2, 2, 2, 2]


def test_level_jumps_are_limited_to_one():
    nb = row_neighbours(6)
    assert list(M.limit_level_jumps(nb, np.array([0, 0, 0, 0, 0, 3], dtype=np.uint8))) == [0, 0, 0, 1, 2, 3]
    assert list(M.limit_level_jumps(nb, np.array([2, 0, 0, 0, 1, 0], dtype=np.uint8))) == [2, 1, 0, 0, 1, 0]


def test_a_landing_takes_equal_steps_with_their_own_levels():
    nb = row_neighbours(8)
    edt = np.array([8.0, 8.0, 8.0, 8.0, 4.0, 2.0, 1.0, 1.0])
    st = M.MultirateStepper(8)
    assert st.assign(edt, nb) == 8.0
    assert list(st.level) == [0, 0, 0, 0, 1, 2, 3, 3]
    dt = st.assign(edt, nb, landing=20.0)
    assert dt == 20.0 / 3.0 and st.stable_step == 8.0
    assert list(st.level) == [0, 0, 0, 0, 1, 2, 3, 3]
    dt = st.assign(edt, nb, landing=3.0)
    assert dt == 3.0 and list(st.level) == [0, 0, 0, 0, 0, 1, 2, 2] and st.finest == 2
    assert st.assign(edt, nb, landing=8.0) == 8.0
    assert list(st.level) == [0, 0, 0, 0, 1, 2, 3, 3]


def test_element_stage_times_and_weights():
    # multirate.rs:169-200: level 1 of a finest-2 step, substep 3, the last SSP-RK3 stage
    w = M.STAGES[M.SSP_RK3][2]
    es = M.element_stage((10.0, 4.0), 2, 3, 2, w, True, 1, 2)
    assert es.time == 10.0 + 1 * 2.0 + 0.5 * 2.0 and es.c == (2.0 / 3.0) * 2.0
    assert es.accumulate == (0.5, 1.0)         # two blocks of the finer rate; not the substep's first
    es = M.element_stage((10.0, 4.0), 2, 2, 1, w, True, 1, 2)
    assert es.accumulate == (0.5, 0.0)         # a level-1 substep starts at depth 1


def test_schedule_visits_every_element_stage_at_its_rate():
    # every element's stage s is evaluated 2^rate[s] times per coarse step, and
    # its substeps end 2^level times
    rng = np.random.default_rng(3)
    mesh = cases.structured_mesh(0.0, 8.0, 0.0, 4.0, 8, 4)
    nb = M.neighbours(mesh.links)
    edt = 2.0 ** -rng.integers(0, 4, size=32).astype(float)
    st = M.MultirateStepper(8)
    st.assign(edt, nb)
    counts = np.zeros((len(st.stages), 32), dtype=int)
    finishes = np.zeros(32, dtype=int)

    def visit(kind, s, elements, plan):
        if kind == "stage":
            counts[s, elements] += 1
        else:
            finishes[elements] += 1

    st.schedule(0.0, 1.0, visit)
    for s in range(len(st.stages)):
        assert np.array_equal(counts[s], 1 << st.rate[s].astype(int))
    assert np.array_equal(finishes, 1 << st.level.astype(int))
    assert st.evaluations_per_step == counts.sum()


# ---------------------------------------------------------------------------
# Whole runs of the restatement
# ---------------------------------------------------------------------------
def oracle_for(c, friction=False, wet_dry=True):
    return O.Oracle.from_case(c, friction=friction, wet_dry=wet_dry)


@pytest.mark.parametrize("scheme", [0, 1])
def test_one_level_is_the_base_scheme_bit_for_bit(scheme):
    c = cases.drying_beach(nx=10, ny=3)
    orc = O.Oracle.from_case(c)
    nb = M.neighbours(c.mesh.links)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    y = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    st = M.MultirateStepper(0, scheme)
    t = 0.0
    for _ in range(20):
        dt = st.assign_levels(orc, nb, x, c.cfl)
        assert st.finest == 0
        st.step(orc, x, t, dt)
        orc.run_dts(*y, t, [dt], scheme)
        t += dt
    for a, b in zip(x, y):
        assert np.array_equal(a, b)


@pytest.fixture(scope="module")
def graded_runs():
    out = {}
    for scheme in (0, 1):
        c = cases.graded_channel(scheme=scheme)
        orc = oracle_for(c, wet_dry=False)
        nb = M.neighbours(c.mesh.links)
        x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
        m = cases.node_masses(c)
        mass0 = float(m @ x[0])
        t, n, st = M.run(orc, nb, x, 0.0, 400.0, 0.5, 8, scheme)
        out[scheme] = (c, x, mass0, t, n, st)
    return out


@pytest.mark.parametrize("scheme", [0, 1])
def test_several_levels_conserve_mass(graded_runs, scheme):
    c, x, mass0, t, n, st = graded_runs[scheme]
    assert t == 400.0
    assert st.stats.finest_level >= 2
    assert st.stats.speedup() > 1.3
    drift = abs(float(cases.node_masses(c) @ x[0]) - mass0) / mass0
    assert drift < 1e-13, drift
    assert all(np.isfinite(a).all() for a in x)
    h = x[0]
    wet = h > 1e-3
    assert (np.hypot(x[1][wet], x[2][wet]) / h[wet]).max() > 1e-3   # the wave crossed the level interfaces


@pytest.mark.parametrize("scheme", [0, 1])
def test_lake_at_rest_with_dry_land_across_levels(scheme):
    c = cases.graded_channel(island=True, friction=True, scheme=scheme, cfl=1.0)
    assert (c.h == 0.0).any(), "the island must be dry"
    orc = oracle_for(c, friction=True, wet_dry=True)
    nb = M.neighbours(c.mesh.links)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    clips0 = orc.clips()
    t, n, st = M.run(orc, nb, x, 0.0, 600.0, 1.0, 8, scheme)
    assert st.stats.finest_level >= 2
    assert orc.clips() == clips0
    assert np.abs(x[0] - c.h).max() < 1e-10
    wet = x[0] > 1e-3
    assert (np.hypot(x[1][wet], x[2][wet]) / x[0][wet]).max(initial=0.0) < 1e-10


# ---------------------------------------------------------------------------
# The product's LTS kernels, emulated on the CPU, against the restatement
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scheme", [0, 1])
def test_emulated_lts_passes_equal_the_oracle(scheme):
    """Every pass of a coarse step with levels 0-3 on the drying beach (wet/dry,
    subcells, friction, tide): the emulated wet_lts / subcell LTS / finish kernels
    give the restatement's stage_where / finish_where bit for bit, listed
    elements and the others (left untouched) alike."""
    from emu_bridge import EmulatedKernels
    c = cases.drying_beach(nx=12, ny=4)
    orc = O.Oracle.from_case(c)
    emu = EmulatedKernels(c)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    orc.run(*x, 0.0, 40, c.cfl)            # a wet/dry state with flow and the tide running
    t0 = 40 * 0.8
    nb = M.neighbours(c.mesh.links)
    rng = np.random.default_rng(7 + scheme)
    edt = 2.0 ** -rng.integers(0, 4, size=c.n_elements).astype(float)
    st = M.MultirateStepper(8, scheme)
    dt = st.assign(edt, nb)
    assert st.finest == 3
    dt = 1.6
    ns = len(st.stages)
    ov = [orc.clone_state(x) for _ in range(ns - 1)]
    ev = [orc.clone_state(x) for _ in range(ns - 1)]
    olast, elast = orc.clone_state(x), orc.clone_state(x)
    oacc, eacc = orc.clone_state(x), orc.clone_state(x)
    ostate, estate = orc.clone_state(x), orc.clone_state(x)
    orc.begin_coarse_step(st.level, st.rate)
    n_sub = 0

    def visit(kind, s, elements, plan):
        nonlocal n_sub
        if kind == "stage":
            oin = ostate if s == 0 else ov[s - 1]
            ein = estate if s == 0 else ev[s - 1]
            last = s + 1 == ns
            orc.stage_where(ostate, oin, elements, plan, olast if last else ov[s], oacc if last else None)
            emu.stage_where(estate, ein, elast if last else ev[s], eacc if last else None, elements, plan)
            oo, eo = (oacc, eacc) if last else (ov[s], ev[s])
            for a, b in zip(oo, eo):
                assert np.array_equal(a, b), (kind, s, plan.substep)
        else:
            orc.finish_where(ostate, oacc, elements, s)
            emu.finish_where(estate, eacc, elements)
            for a, b in zip(ostate, estate):
                assert np.array_equal(a, b), (kind, s)
        _, es = orc.masks(ostate[0])
        n_sub += int(es.sum())

    st.schedule(t0, dt, visit)
    assert n_sub > 0, "no subcell element: the subcell LTS kernel went untested"


def test_emulated_forced_lts_passes_equal_the_oracle():
    """The forced instantiations of the LTS kernels (wet_lts_kernel<N, true>, the
    subcell kernel's), emulated, under tides per boundary face node with a velocity
    and transport scaling (11 constituents: each level's moment row at its own stage
    time), Manning n per node and a beta-plane: every pass of a coarse step with
    levels 0-3 on the drying beach equals the restatement bit for bit."""
    from emu_bridge import EmulatedKernels
    from dgcom_amd import tides as T
    c = cases.drying_beach(nx=12, ny=4)
    rng = np.random.default_rng(11)
    omega = 2.0 * np.pi / np.array([T.constituent_period(n) for n in ("M2", "S2", "N2", "K1", "O1", "2N2", "MU2",
                                                                      "NU2", "LDA2", "L2", "T2")])
    a = rng.uniform(0.01, 0.2, size=(omega.size, 6))
    bt = cases.boundary_tides(c, omega, rng.uniform(-1e-9, 1e-9, omega.size),
                              lambda x, y: a * (1.0 + 0.2 * np.sin(y / 120.0)), ramp=600.0, velocity=True,
                              max_ratio=1.5, mean=lambda x, y: 0.02 * np.cos(y / 150.0),
                              source_depth=lambda x, y: 4.0 + 0.002 * y)
    mn = np.repeat(c.manning, c.nn) * rng.uniform(0.6, 1.4, c.nn * c.n_elements)
    orc = O.Oracle.from_case(c)
    orc.set_boundary_tides(bt)
    orc.set_manning_nodal(mn)
    orc.set_beta(2e-8, 300.0)
    emu = EmulatedKernels(c)
    cf = emu.coriolis_nodes(2e-8, 300.0)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    orc.run(*x, 0.0, 40, c.cfl)
    t0 = 40 * 0.8
    nb = M.neighbours(c.mesh.links)
    edt = 2.0 ** -np.random.default_rng(3).integers(0, 4, size=c.n_elements).astype(float)
    st = M.MultirateStepper(8, 1)
    st.assign(edt, nb)
    assert st.finest == 3
    ns = len(st.stages)
    ov = [orc.clone_state(x) for _ in range(ns - 1)]
    ev = [orc.clone_state(x) for _ in range(ns - 1)]
    olast, elast = orc.clone_state(x), orc.clone_state(x)
    oacc, eacc = orc.clone_state(x), orc.clone_state(x)
    ostate, estate = orc.clone_state(x), orc.clone_state(x)
    orc.begin_coarse_step(st.level, st.rate)
    n_stage = 0

    def visit(kind, s, elements, plan):
        nonlocal n_stage
        if kind == "stage":
            oin = ostate if s == 0 else ov[s - 1]
            ein = estate if s == 0 else ev[s - 1]
            last = s + 1 == ns
            orc.stage_where(ostate, oin, elements, plan, olast if last else ov[s], oacc if last else None)
            emu.stage_where_forced(estate, ein, elast if last else ev[s], eacc if last else None, elements, plan,
                                   bt, mn, cf)
            oo, eo = (oacc, eacc) if last else (ov[s], ev[s])
            for a_, b_ in zip(oo, eo):
                assert np.array_equal(a_, b_), (kind, s, plan.substep)
            n_stage += 1
        else:
            orc.finish_where(ostate, oacc, elements, s)
            emu.finish_where(estate, eacc, elements)
            for a_, b_ in zip(ostate, estate):
                assert np.array_equal(a_, b_), (kind, s)

    st.schedule(t0, 1.6, visit)
    assert n_stage > 0


def test_line_local_rates_cross_an_element_in_two_stages():
    # multirate.rs:1511-1523: a row of line-local elements, the last on level 1:
    # the finer rate reaches one element further every two stages
    n = 6
    mesh = cases.structured_mesh(0.0, float(n), 0.0, 1.0, n, 1)
    nb = M.neighbours(mesh.links)
    n1 = 3
    level = np.array([0, 0, 0, 0, 0, 1], dtype=np.uint8)
    frm = np.repeat(level[:, None], n1 * n1, axis=1)
    ll = np.ones(n, dtype=bool)
    rates = []
    for _ in range(4):
        frm, rate = M.spread_over_lines(nb, mesh.links, n1, ll, frm)
        rates.append(list(rate))
    assert rates == [[0, 0, 0, 0, 1, 1], [0, 0, 0, 0, 1, 1], [0, 0, 0, 1, 1, 1], [0, 0, 0, 1, 1, 1]]
    # whole elements: one face per stage
    frm = np.repeat(level[:, None], n1 * n1, axis=1)
    rates = []
    for _ in range(2):
        frm, rate = M.spread_over_lines(nb, mesh.links, n1, np.zeros(n, dtype=bool), frm)
        rates.append(list(rate))
    assert rates == [[0, 0, 0, 0, 1, 1], [0, 0, 0, 1, 1, 1]]


@pytest.mark.parametrize("scheme", [0, 1])
def test_line_local_rates_are_the_whole_element_rates_bit_for_bit(scheme):
    """local_time_stepping_test.rs:1122-1200: a run on node-by-node rates is the
    run on whole-element rates bit for bit, with fewer evaluations."""
    c = cases.graded_channel(scheme=scheme)
    nb = M.neighbours(c.mesh.links)
    out = []
    for line_rates in (True, False):
        orc = oracle_for(c, wet_dry=False)
        x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
        st = M.MultirateStepper(8, scheme, line_rates=line_rates)
        t, n, st = M.run(orc, nb, x, 0.0, 200.0, 0.5, 8, scheme, stepper=st)
        out.append((x, n, st.stats))
    (xa, na, sa), (xb, nb_, sb) = out
    assert na == nb_
    for a, b in zip(xa, xb):
        assert np.array_equal(a, b)
    assert sa.element_evaluations < sb.element_evaluations


# ---------------------------------------------------------------------------
# The reference's published local-time-stepping known answers (notes/gpu.md:184)
# ---------------------------------------------------------------------------
LTS_KNOWN_STEPS = {"at_rest": 1470, "incoming_wave": 1529}


@pytest.mark.parametrize("velocity", ["at_rest", "incoming_wave"])
@pytest.mark.parametrize("max_levels", [3, 8])
def test_tidal_channel_coarse_steps(velocity, max_levels):
    """cases/tidal_channel over 25 h with local time stepping, landing on the
    30-minute outputs: 1,470 coarse steps (1,529 with the incoming wave), at 3
    levels and at 8 (the channel uses 3), no clip; the mass conserved."""
    c = cases.tidal_channel(velocity=velocity)
    orc = O.Oracle.from_case(c)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    nb = M.neighbours(c.mesh.links)
    t, n, st = M.run(orc, nb, x, 0.0, 25 * 3600.0, c.cfl, max_levels, c.scheme, interval=1800.0)
    assert n == LTS_KNOWN_STEPS[velocity]
    assert st.stats.finest_level == 2 and orc.clips() == 0


def test_refined_tidal_channel_levels():
    """The channel refined to 64,000 elements, 15 min at 8 levels: 143 coarse
    steps of 6.29 s with 2^3 finest substeps, 1.23x fewer element evaluations
    than global steps (notes/gpu.md:184,186)."""
    c = cases.tidal_channel(nx=800, ny=80)
    orc = O.Oracle.from_case(c)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    nb = M.neighbours(c.mesh.links)
    st = M.MultirateStepper(8, c.scheme)
    dt = st.assign_levels(orc, nb, x, c.cfl, None, 900.0)
    assert dt == 900.0 / 143 and st.finest == 3
    assert abs(st.planned_speedup() - 1.23) < 0.005


# ---------------------------------------------------------------------------
# The library's own stepper (C++, multirate::Assigner) against this one
# ---------------------------------------------------------------------------
class _Bound:
    def __init__(self, wet_dry):
        self.positivity_h_dry = 1e-3 if wet_dry else None


def _py_assign(st, edt, flags, nb, links, n1, wet_dry, cap, landing, line_rates):
    bound = M.MultirateStepper.coarse_step_bound(_Bound(wet_dry), edt, flags)
    c = bound if cap is None else min(cap, bound)
    lines = (n1, links, (flags & 4) != 0) if line_rates else None
    dt = st.assign(edt, nb, c if math.isfinite(c) else None, landing, lines)
    return dt, st.stable_step, st.finest, st.level.copy(), np.stack(st.rate)


@pytest.mark.parametrize("line_rates", [True, False])
@pytest.mark.parametrize("scheme", [0, 1])
def test_library_assigner_equals_the_stepper(line_rates, scheme):
    """Levels, coarse steps, the limit and the stage rates of the library's C++
    stepper equal this module's over a sequence of calls on a jittered mesh:
    random element steps (some infinite), depth flags, caps and landings, calls
    repeated with the same inputs (the kept rates) and with changed ones."""
    from emu_bridge import multirate_assign
    mesh = cases.structured_mesh(0.0, 1.0, 0.0, 1.0, 13, 9, jitter=0.2, seed=5)
    K = mesh.n_elements
    nb = M.neighbours(mesh.links)
    rng = np.random.default_rng(11 + scheme)
    n1 = 3
    calls = []
    for q in range(10):
        edt = 2.0 ** rng.uniform(-4.0, 0.0, size=K)
        edt[rng.random(K) < 0.05] = math.inf
        flags = rng.integers(0, 8, size=K).astype(np.uint8)
        cap = [None, 0.7, 0.2][q % 3]
        landing = [None, 3.3, 0.9, 1e-3][q % 4]
        calls.append((edt, flags, cap, landing))
        if q % 3 == 0:
            calls.append((edt, flags, cap, landing))     # the same again: the kept rates
    for wet_dry in (True, False):
        ns = len(M.STAGES[scheme])
        got = multirate_assign(mesh.links, n1, 8, ns, line_rates, [c[0] for c in calls], [c[1] for c in calls],
                               wet_dry, [c[2] for c in calls], [c[3] for c in calls])
        st = M.MultirateStepper(8, scheme, line_rates=line_rates)
        for (edt, flags, cap, landing), g in zip(calls, got):
            want = _py_assign(st, edt, flags, nb, mesh.links, n1, wet_dry, cap, landing, line_rates)
            assert g[0] == want[0] and g[1] == want[1] and g[2] == want[2]
            assert np.array_equal(g[3], want[3])
            assert np.array_equal(g[4], want[4])
        # the library's device half (extremes, plan, levels and the comparison
        # with the kept assignment, emulated) gives the same sequence
        dev = multirate_assign(mesh.links, n1, 8, ns, line_rates, [c[0] for c in calls], [c[1] for c in calls],
                               wet_dry, [c[2] for c in calls], [c[3] for c in calls], device=True)
        for g, d in zip(got, dev):
            assert g[0] == d[0] and g[1] == d[1] and g[2] == d[2]
            assert np.array_equal(g[3], d[3]) and np.array_equal(g[4], d[4])


def test_library_assigner_on_the_refined_tidal_channel():
    from emu_bridge import multirate_assign
    c = cases.tidal_channel(nx=800, ny=80)
    orc = O.Oracle.from_case(c)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    edt = orc.element_dt_state(x, c.cfl, c.scheme)
    flags = orc.depth_flags(x, 0.1, 1e-3, 1e-3 + 1.0)
    (dt, stable, finest, level, rates), = multirate_assign(c.mesh.links, 3, 8, 4, True, [edt], [flags], True, [None],
                                                           [900.0])
    assert dt == 900.0 / 143 and finest == 3
    nb = M.neighbours(c.mesh.links)
    st = M.MultirateStepper(8, c.scheme)
    assert st.assign_levels(orc, nb, x, c.cfl, None, 900.0) == dt
    assert np.array_equal(st.level, level) and np.array_equal(np.stack(st.rate), rates)
    # through the device half (emulated) and the host remainder on its threads (64,000 elements: two parts)
    (d_dt, d_stable, d_finest, d_level, d_rates), = multirate_assign(c.mesh.links, 3, 8, 4, True, [edt], [flags], True,
                                                                     [None], [900.0], device=True)
    assert (d_dt, d_stable, d_finest) == (dt, stable, finest)
    assert np.array_equal(d_level, level) and np.array_equal(d_rates, rates)
