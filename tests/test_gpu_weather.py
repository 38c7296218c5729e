"""Gridded weather on the GPU against the restatement: whole adaptive runs bit
for bit under wind and pressure fields varying in space and time across the
drag laws' branches (Large & Pond with both switches and linear snapshots; Wu,
the wind alone, cubic snapshots and a ramp; the pressure alone), the open
boundary raised by the inverse barometer (on the uniform tide and on tides per
node), and a stage
time the snapshots do not cover, the ensembles and the VJP refused by name (local
time stepping: tests/test_gpu_forcing_lts.py)."""
import numpy as np
import pytest

import oracle as O
from dgcom_amd import cases, swe2d

pytestmark = pytest.mark.gpu

T0 = 1_706_594_400.0


def _weather(case, **options):
    times = T0 + np.arange(-1800.0, 4 * 3600.0, 1800.0)

    def fields(x, y, t):
        s = (t - T0) / 2000.0
        u = 8.0 + 20.0 * np.sin(x / 300.0 + s)          # through 11 and 25 m/s
        v = 6.0 * np.cos(y / 200.0 - s)
        return u, v, 0.002 * np.sin(x / 400.0 + s), 0.001 * np.cos(y / 300.0), np.zeros_like(x)

    return cases.weather_at_nodes(case, times, fields, epoch=T0, **options)


VARIANTS = [
    ("large_pond", dict()),
    ("wu_cubic_ramp", dict(drag=cases.DRAG_WU, pressure=False, cubic=True, ramp=40.0)),
    ("pressure", dict(wind=False, h_min=0.05)),
]


@pytest.mark.parametrize("name,options", VARIANTS, ids=[v[0] for v in VARIANTS])
def test_runs_equal_the_oracle(name, options):
    c = cases.drying_beach(nx=16, ny=6, jitter=0.15)
    wx = _weather(c, **options)
    orc = O.Oracle.from_case(c)
    orc.set_weather(wx)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_weather(wx)
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


@pytest.mark.parametrize("factor,atlas", [(1.0, False), (0.25, True)], ids=["uniform_tide", "tides_per_node"])
def test_the_inverse_barometer_equals_the_oracle(factor, atlas):
    # the open boundary raised by the weather's inverse barometer over a pressure field varying along it
    c = cases.drying_beach(nx=16, ny=6, jitter=0.15)
    n = c.h.size
    times = T0 + np.arange(0.0, 4 * 3600.0, 1800.0)
    wx = cases.weather_at_nodes(c, times, lambda x, y, t: (np.full(n, 5.0), np.zeros(n), 0.002 * np.ones(n),
                                                           np.zeros(n), 800.0 * np.sin(y / 150.0 + (t - T0) / 900.0)),
                                epoch=T0, ramp=30.0)
    orc = O.Oracle.from_case(c)
    phys = swe2d.SWEPhysics2D.from_case(c)
    for m in (orc, phys):
        m.set_weather(wx)
        m.set_inverse_barometer(factor)
    if atlas:
        bt = cases.boundary_tides(c, [2 * np.pi / 600.0], [0.0], lambda x, y: np.array([[0.3, 0.1, 0.2, 0.0, 0.05, 0.0]]),
                                  velocity=True, mean=lambda x, y: 0.02)
        orc.set_boundary_tides(bt)
        phys.set_boundary_tides(bt)
    s = phys.state(c.h, c.hu, c.hv)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    t_g = phys.run(s, 0.0, 60, c.cfl)
    t_o = orc.run(*x, 0.0, 60, c.cfl)
    assert t_g == t_o
    for g, r in zip(s.download(), x):
        assert np.array_equal(g, r)
    ref = O.Oracle.from_case(c)
    ref.set_weather(wx)
    if atlas:
        ref.set_boundary_tides(bt)
    y = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    ref.run(*y, 0.0, 60, c.cfl)
    assert np.abs(y[0] - x[0]).max() > 1e-6          # the raise acts


def test_refusals():
    c = cases.drying_beach(nx=8, ny=3)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_weather(_weather(c))
    s = phys.state(c.h, c.hu, c.hv)
    with pytest.raises(RuntimeError, match="do not cover"):
        phys.step(s, 5.0 * 3600.0, 0.1)
    w, gb, gi = phys.state(c.h, c.hu, c.hv), phys.state().zero(), phys.state().zero()
    with pytest.raises(RuntimeError, match="no VJP"):
        phys.stage_vjp(s, s, 1.0, 0.0, 1.0, 0.0, w, gb, gi)
    with pytest.raises(RuntimeError, match="weather"):
        swe2d.Ensemble([phys])
    phys.set_weather(None)
    phys.step(s, 5.0 * 3600.0, 0.1)
