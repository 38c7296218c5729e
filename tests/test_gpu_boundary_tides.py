"""Tides per boundary node on the GPU against the restatement: whole adaptive
runs bit for bit with per-node constants varying along the open boundary (with
a velocity, its transport scaling and a mean level; and elevation only as an
incoming wave), the tidal channel with its tide written as node constants
within rounding of the uniform tide's run; the VJP refuses them by name (local
time stepping: tests/test_gpu_forcing_lts.py)."""
import numpy as np
import pytest

import oracle as O
from dgcom_amd import cases, swe2d

pytestmark = pytest.mark.gpu


def _varying(case, velocity, ratio):
    rng = np.random.default_rng(7)
    a = rng.uniform(0.2, 0.6, size=(2, 6))

    def constants(x, y):
        w = 1.0 + 0.3 * np.sin(y / 150.0)
        return a * w * np.array([1.0, 1.0, 0.3, 0.3, 0.1, 0.1])

    return cases.boundary_tides(case, [2 * np.pi / 600.0, 2 * np.pi / 400.0], [1e-7, -2e-7], constants,
                                ramp=120.0, velocity=velocity, max_ratio=ratio,
                                mean=lambda x, y: 0.05 * np.cos(y / 200.0),
                                source_depth=lambda x, y: 6.0 + 0.001 * y)


@pytest.mark.parametrize("velocity,ratio,incoming", [(True, 2.0, False), (True, 0.0, False), (False, 0.0, True)])
def test_runs_equal_the_oracle(velocity, ratio, incoming):
    c = cases.drying_beach(nx=16, ny=6, jitter=0.15)
    c.incoming_wave = incoming
    bt = _varying(c, velocity, ratio)
    orc = O.Oracle.from_case(c)
    orc.set_boundary_tides(bt)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_boundary_tides(bt)
    s = phys.state(c.h, c.hu, c.hv)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    t_g = phys.run(s, 0.0, 60, c.cfl)
    t_o = orc.run(*x, 0.0, 60, c.cfl)
    assert t_g == t_o
    for g, r in zip(s.download(), x):
        assert np.array_equal(g, r)
    assert phys.negative_depth_clips() == orc.clips()


def test_the_tidal_channel_with_node_constants():
    c = cases.tidal_channel()
    phys_u = swe2d.SWEPhysics2D.from_case(c)
    phys_a = swe2d.SWEPhysics2D.from_case(c)
    phys_a.set_boundary_tides(cases.boundary_tides_of(c, c.tide))
    su, sa = phys_u.state(c.h, c.hu, c.hv), phys_a.state(c.h, c.hu, c.hv)
    dts = []
    t = 0.0
    for _ in range(400):
        dt = phys_u.compute_dt_ssp(su, c.cfl, c.scheme)
        phys_u.step(su, t, dt, c.scheme)
        phys_a.step(sa, t, dt, c.scheme)
        t += dt
    hu_, ha = su.download()[0], sa.download()[0]
    assert np.abs(hu_ - ha).max() < 1e-9


def test_refusals():
    c = cases.drying_beach(nx=8, ny=3)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_boundary_tides(_varying(c, True, 0.0))
    s = phys.state(c.h, c.hu, c.hv)
    w, gb, gi = phys.state(c.h, c.hu, c.hv), phys.state().zero(), phys.state().zero()
    with pytest.raises(RuntimeError, match="no VJP"):
        phys.stage_vjp(s, s, 1.0, 0.0, 1.0, 0.0, w, gb, gi)
    phys.set_boundary_tides(None)
    phys.step(s, 0.0, 0.1)
