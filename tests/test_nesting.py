"""The nesting band (NestingRelaxation2D, src/boundary/ocean_nesting.rs) in the
restatement, pinned by what the mathematics fixes: in a closed basin at rest
under a uniform parent the RHS is the relaxation gamma (eta_p - eta) alone, and
SSP-RK3 follows the exponential decay to the parent; snapshot interpolation
is exact for a parent linear in time (linear stencil) and cubic in time (cubic
stencil), the linear stencil not for the cubic; the ramp's smoothstep; momentum
only keeps the mass; a parent velocity scaled by D_parent / D_child, limited;
no source on dry nodes or outside the band. The GPU build is held to the
restatement in tests/test_gpu_nesting.py."""
import math

import numpy as np
import pytest

import oracle as O
from dgcom_amd import cases

G = 9.81
T0 = 1_706_594_400.0       # the Unix epoch of model time 0 in the reference's nesting tests


def _basin(depth=5.0, n=4, order=2):
    mesh = cases.structured_mesh(0.0, 400.0, 0.0, 400.0, n, n, sides=(cases.WALL, cases.WALL, cases.WALL, cases.TIDE))
    x, _ = cases.node_coordinates(mesh, order)
    bed = np.full(x.size, -depth)
    h, hu, hv = cases.still_water(bed)
    return cases.Case("basin", order, mesh, bed, np.zeros(mesh.n_elements), 0.0, cases.Tide(), h, hu, hv)


def _closed(case):
    # every face a wall: the band is set over every node by hand
    case.mesh.tags[:] = np.where(case.mesh.links == cases.BOUNDARY, cases.WALL, case.mesh.tags)
    return case


def _uniform(case, times, zeta, u=0.0, v=0.0, **kw):
    n = case.h.size
    series = np.empty((n, len(times), 3), dtype=np.float32)
    for j, t in enumerate(times):
        series[:, j] = (zeta(t), u, v)
    return cases.Nesting(np.arange(n, dtype=np.uint32), np.asarray(times, dtype=np.float64), series, np.ones(n),
                         np.full(n, np.nan), **kw)


def test_uniform_relaxation_in_a_closed_basin():
    c = _closed(_basin())
    ns = _uniform(c, [T0], lambda t: 0.25, timescale=600.0)
    orc = O.Oracle.from_case(c, friction=False)
    orc.set_nesting(ns)
    h, hu, hv = orc.rhs(c.h, c.hu, c.hv, 0.0)
    assert np.allclose(h, 0.25 / 600.0, rtol=1e-13, atol=0.0)
    assert np.abs(hu).max() < 1e-14 and np.abs(hv).max() < 1e-14
    # the decay eta(t) = 0.25 (1 - exp(-t / tau)), integrated by SSP-RK3
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    dt, n = 2.0, 300       # within the basin's CFL bound
    orc.run_dts(*x, 0.0, [dt] * n)
    expect = 5.0 + 0.25 * (1.0 - math.exp(-dt * n / 600.0))
    assert np.abs(x[0] - expect).max() < 1e-6
    assert np.ptp(x[0]) < 1e-12


@pytest.mark.parametrize("cubic", [False, True])
def test_snapshot_interpolation(cubic):
    c = _closed(_basin())
    times = T0 + np.array([0.0, 3600.0, 7200.0, 10800.0, 14400.0])
    lin = lambda t: 0.1 + 0.2 * (t - T0) / 14400.0                       # noqa: E731
    cub = lambda t: 0.1 + 0.3 * ((t - T0) / 14400.0) ** 3                # noqa: E731
    # the cubic stencil is quadratic in the first and last interval (0-3600 s, 10800-14400 s)
    for zeta, exact_at in ((lin, lambda t: True), (cub, lambda t: t in (0.0, 3600.0) or (cubic and 3600.0 <= t <= 10800.0))):
        ns = _uniform(c, times, zeta, timescale=100.0, epoch=T0, cubic=cubic)
        orc = O.Oracle.from_case(c, friction=False)
        orc.set_nesting(ns)
        for t in (0.0, 1000.0, 3600.0, 5400.0, 9000.0, 14000.0):
            got = orc.rhs(c.h, c.hu, c.hv, t)[0] * 100.0
            # the parent at t from its f32 snapshots: exact up to their rounding
            err = np.abs(got - zeta(T0 + t)).max()
            if exact_at(t):
                assert err < 2e-8, (zeta, t, err)
            else:
                assert err > 1e-5, (zeta, t, err)


def test_the_ramp_and_momentum_only():
    c = _closed(_basin())
    # values exact in f32, so that the closed forms hold to rounding
    ns = _uniform(c, [T0], lambda t: 0.375, u=0.25, v=-0.125, timescale=50.0, ramp=1000.0, velocity=True)
    orc = O.Oracle.from_case(c, friction=False)
    orc.set_nesting(ns)
    for t, r in ((0.0, 0.0), (250.0, 0.15625), (500.0, 0.5), (1000.0, 1.0), (5000.0, 1.0)):
        h, hu, hv = orc.rhs(c.h, c.hu, c.hv, t)
        d = 5.0 + r * 0.375
        assert np.allclose(h, r * 0.375 / 50.0, rtol=1e-12, atol=1e-13)
        assert np.allclose(hu, d * r * 0.25 / 50.0, rtol=1e-12, atol=1e-13)
        assert np.allclose(hv, -d * r * 0.125 / 50.0, rtol=1e-12, atol=1e-13)
    ns.momentum_only = True
    orc.set_nesting(ns)
    h, hu, hv = orc.rhs(c.h, c.hu, c.hv, 500.0)
    assert np.abs(h).max() < 1e-14
    assert np.allclose(hu, 5.0 * 0.5 * 0.25 / 50.0, rtol=1e-12, atol=1e-13)


def test_transport_scaling_and_dry_nodes():
    c = _closed(_basin())
    ns = _uniform(c, [T0], lambda t: 0.0, u=0.5, timescale=10.0, velocity=True)
    for ratio, depth, scale in ((2.0, 12.0, 2.0), (3.0, 12.0, 2.4), (3.0, 1.0, 1.0 / 3.0), (0.0, 12.0, 1.0),
                                (3.0, np.nan, 1.0)):
        ns.max_ratio = ratio
        ns.depth = np.full(ns.weight.size, depth)
        orc = O.Oracle.from_case(c, friction=False)
        orc.set_nesting(ns)
        hu = orc.rhs(c.h, c.hu, c.hv, 0.0)[1]
        assert np.allclose(hu, 5.0 * 0.5 * scale / 10.0, rtol=1e-13, atol=1e-13), (ratio, depth)
    # a child shallower than the wet depth, and a parent drier than it, get nothing
    h = c.h.copy()
    h[::3] = 0.01
    ns.max_ratio = 0.0
    orc = O.Oracle.from_case(c, friction=False)
    orc.set_nesting(ns)
    r = orc.rhs(h, c.hu, c.hv, 0.0)
    dry = O.Oracle.from_case(c, friction=False).rhs(h, c.hu, c.hv, 0.0)
    for a, b in zip(r, dry):
        assert np.array_equal(a[::3], b[::3])
    ns2 = _uniform(c, [T0], lambda t: -4.97, timescale=10.0)
    orc.set_nesting(ns2)
    for a, b in zip(orc.rhs(c.h, c.hu, c.hv, 0.0), O.Oracle.from_case(c, friction=False).rhs(c.h, c.hu, c.hv, 0.0)):
        assert np.array_equal(a, b)


def test_the_band():
    # nodes on the open face weigh 1, the cosine profile inside the band, none beyond
    c = _basin(n=8, order=3)
    ns = cases.nesting_band(c, 120.0, [T0], lambda x, y, t: (0.125, 0.0, 0.0), epoch=T0)
    x, y = (a.ravel() for a in cases.node_coordinates(c.mesh, c.order))
    inside = ns.slot != 0xFFFFFFFF
    d = x                  # the open side is the west one (sides: south, east, north, west)
    assert np.array_equal(inside, d < 120.0 - 1e-9) or np.array_equal(inside, d < 120.0)
    w = ns.weight[ns.slot[inside]]
    expect = np.where(d[inside] < 1e-9, 1.0, 0.5 * (1.0 - np.cos(np.pi * (1.0 - d[inside] / 120.0))))
    assert np.allclose(w, expect, rtol=0, atol=1e-12)
    orc = O.Oracle.from_case(c, friction=False)
    orc.set_nesting(ns)
    h = orc.rhs(c.h, c.hu, c.hv, 0.0)[0]
    assert np.abs(h[~inside]).max() < 1e-14
    assert np.allclose(h[inside], w * 0.125 / 3600.0, rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("u_p", [0.0, 0.5])
def test_the_parent_as_the_open_boundary(u_p):
    # the west face's external state is the parent's (level 0.125, velocity u_p along +x): with the interior at
    # rest, w_out = 2 c_i and w_in = -u_p - 2 c_e, so u_n,b = c_i - c_e - u_p / 2 and c_b = (c_i + c_e) / 2 + u_p / 4
    mesh = cases.structured_mesh(0.0, 400.0, 0.0, 100.0, 4, 1, sides=(cases.WALL, cases.WALL, cases.WALL, cases.TIDE))
    x, _ = cases.node_coordinates(mesh, 2)
    bed = np.full(x.size, -5.0)
    h, hu, hv = cases.still_water(bed)
    c = cases.Case("channel", 2, mesh, bed, np.zeros(mesh.n_elements), 0.0, cases.Tide(), h, hu, hv)
    ns = cases.nesting_band(c, 1.0, [T0], lambda x, y, t: (0.125, u_p, 0.0), epoch=T0, velocity=u_p != 0.0)
    ns.weight[:] = 0.0                      # the boundary alone, no relaxation
    orc = O.Oracle.from_case(c, friction=False)
    orc.set_nesting(ns)
    orc.set_nesting_boundary(True)
    fm = orc.face_mass(c.h, c.hu, c.hv, 0.0).ravel()
    west = [(k * 4 + 3) * 3 + fi for k in range(c.n_elements) for fi in range(3) if c.mesh.links[k, 3] == cases.BOUNDARY]
    ci, ce = math.sqrt(G * 5.0), math.sqrt(G * 5.125)
    cb = (ci + ce) / 2.0 + u_p / 4.0
    assert np.allclose(fm[west], cb * cb / G * (ci - ce - u_p / 2.0), rtol=1e-12, atol=1e-15)
    orc.set_nesting_boundary(False)
    assert np.abs(orc.face_mass(c.h, c.hu, c.hv, 0.0).ravel()[west]).max() < 1e-15
