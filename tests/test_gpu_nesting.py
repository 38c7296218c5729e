"""The nesting band on the GPU against the restatement: whole adaptive runs bit
for bit with a band over the open boundary relaxing towards a parent that
varies in space and time (elevation only with linear snapshots; a velocity with
its transport scaling, the datum shift, the ramp and cubic snapshots; momentum
only; the parent as the open boundary's external state too), and a stage time the snapshots do not cover, the ensembles and the VJP
refused by name (local time stepping: tests/test_gpu_forcing_lts.py)."""
import numpy as np
import pytest

import oracle as O
from dgcom_amd import cases, swe2d

pytestmark = pytest.mark.gpu

T0 = 1_706_594_400.0


def _band(case, **options):
    times = T0 + np.arange(-600.0, 4 * 3600.0, 600.0)

    def parent(x, y, t):
        ph = 2.0 * np.pi * (t - T0) / 3600.0 + y / 300.0
        return 0.05 * np.sin(ph), 0.1 * np.cos(ph), 0.02 + 0.01 * np.sin(x / 50.0)

    return cases.nesting_band(case, 200.0, times, parent, parent_depth=lambda x, y: 1.5 + 0.002 * y, epoch=T0,
                              **options)


VARIANTS = [
    ("elevation", dict(timescale=300.0)),
    ("velocity", dict(timescale=120.0, velocity=True, max_ratio=2.0, reference_level=0.03, ramp=50.0, cubic=True)),
    ("momentum", dict(timescale=200.0, velocity=True, momentum_only=True, wet_depth=0.1)),
    ("boundary", dict(timescale=150.0, velocity=True, max_ratio=2.0, cubic=True)),
]


@pytest.mark.parametrize("name,options", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_runs_equal_the_oracle(name, options):
    c = cases.drying_beach(nx=16, ny=6, jitter=0.15)
    ns = _band(c, **options)
    assert (ns.slot != 0xFFFFFFFF).sum() > c.nn * 6          # more than the column of elements at the boundary
    orc = O.Oracle.from_case(c)
    orc.set_nesting(ns)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_nesting(ns)
    if name == "boundary":
        # the open boundary's external state from the parent too
        orc.set_nesting_boundary(True)
        phys.set_nesting_boundary(True)
    s = phys.state(c.h, c.hu, c.hv)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    t_g = phys.run(s, 0.0, 60, c.cfl)
    t_o = orc.run(*x, 0.0, 60, c.cfl)
    assert t_g == t_o
    for g, r in zip(s.download(), x):
        assert np.array_equal(g, r)
    assert phys.negative_depth_clips() == orc.clips()
    out = phys.state()
    phys.compute_rhs_into(s, t_g, out)
    for a, b in zip(out.download(), orc.rhs(*x, t_g)):
        assert np.array_equal(a, b)


def test_refusals():
    c = cases.drying_beach(nx=8, ny=3)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_nesting(_band(c, timescale=300.0))
    s = phys.state(c.h, c.hu, c.hv)
    with pytest.raises(RuntimeError, match="do not cover"):
        phys.step(s, 5.0 * 3600.0, 0.1)
    w, gb, gi = phys.state(c.h, c.hu, c.hv), phys.state().zero(), phys.state().zero()
    with pytest.raises(RuntimeError, match="no VJP"):
        phys.stage_vjp(s, s, 1.0, 0.0, 1.0, 0.0, w, gb, gi)
    with pytest.raises(RuntimeError, match="nesting"):
        swe2d.Ensemble([phys])
    phys.set_nesting(None)
    phys.step(s, 5.0 * 3600.0, 0.1)
    # the parent as the open boundary's state needs a slot at every open face node
    with pytest.raises(RuntimeError, match="no parent"):
        phys.set_nesting_boundary(True)
    ns = _band(c, timescale=300.0)
    n1 = c.order + 1
    k, f, fi, _, _ = cases.open_face_nodes(c)[0]
    ns.slot[k * c.nn + cases.face_node(n1, f, fi)] = 0xFFFFFFFF
    phys.set_nesting(ns)
    with pytest.raises(RuntimeError, match=f"open face node {fi} of face {f} of element {k} has no parent slot"):
        phys.set_nesting_boundary(True)


# ---------------------------------------------------------------------------
# The reference's nesting gate (gpu/src/nest.rs, notes/gpu.md:199): the GPU equals the
# restatement bit for bit over its 8 h, with global steps and with 3 levels
# ---------------------------------------------------------------------------
GATE = {"cubic": (10845, 2285), "linear": (10862, 2293)}


def _landed_steps(step_dt, t_end):
    t, n = 0.0, 0
    while t < t_end:
        dt = step_dt(t, None)
        landed = t + dt * (1.0 + 1e-9) >= t_end
        if landed:
            dt = t_end - t
        step_dt(t, dt)
        t = t_end if landed else t + dt
        n += 1
    return n


@pytest.mark.parametrize("variant", ["cubic", "linear"])
def test_the_nesting_gate_equals_the_restatement(variant):
    from dgcom_amd import multirate as M, nesting as NS
    c, pm = NS.gate_case(variant)
    t_end = NS.GATE_HOURS * 3600.0
    orc = O.Oracle.from_case(c)
    orc.set_nesting(pm.nesting)
    orc.set_nesting_boundary(True)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]

    def orc_step(t, dt):
        if dt is None:
            return orc.dt(*x, c.cfl, c.scheme)
        orc.run_dts(*x, t, [dt], c.scheme)

    n_o = _landed_steps(orc_step, t_end)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_nesting(pm.nesting)
    phys.set_nesting_boundary(True)
    s = phys.state(c.h, c.hu, c.hv)

    def gpu_step(t, dt):
        if dt is None:
            return phys.compute_dt_ssp(s, c.cfl, c.scheme)
        phys.step(s, t, dt, c.scheme)

    n_g = _landed_steps(gpu_step, t_end)
    assert n_g == n_o == GATE[variant][0]
    for a, b in zip(s.download(), x):
        assert np.array_equal(a, b)
    assert phys.negative_depth_clips() == orc.clips() == 0
    # three levels of local time stepping
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    orc = O.Oracle.from_case(c)
    orc.set_nesting(pm.nesting)
    orc.set_nesting_boundary(True)
    nb = M.neighbours(c.mesh.links)
    _, n_o, _ = M.run(orc, nb, x, 0.0, t_end, c.cfl, 3, c.scheme)
    s2 = phys.state(c.h, c.hu, c.hv)
    local = phys.local_time_stepping()
    _, n_g, st = M.run(local, nb, s2, 0.0, t_end, c.cfl, 3, c.scheme)
    st.release(local)
    assert n_g == n_o == GATE[variant][1]
    for a, b in zip(s2.download(), x):
        assert np.array_equal(a, b)
