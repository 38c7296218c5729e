"""The reference's nesting gate (gpu/src/nest.rs, notes/gpu.md:196-199) on the
restatement: the tidal channel at 500 m (640 P2 elements, 48 in a 3 km band)
nested in its made-up parent (an M2-like wave entering the channel, a
quarter-diurnal, a coastal current, hourly on an 800 m grid, its depth 32-42 m),
the parent sampled at the open boundary and the band (dgcom_amd.nesting), the bed
blended to the parent's, Coriolis, Manning, the wet/dry stage, 8 h:

- cubic in time and transport-scaled, relaxed at 30 min: 10,845 steps, and 2,285
  coarse steps at 3 levels of local time stepping;
- linear, unscaled, the momentum alone relaxed at 15 min: 10,862 steps, and 2,293
  coarse steps at 3 levels.

Also the sampler's pieces against the reference's unit tests: the band's
distances (band.rs:125-160), the bilinear stencil over wet corners
(geo_grid.rs:420-450), the geodesic distance of a tenth of a degree
(geo_grid.rs:844), east in a local projection (geo_grid.rs:807). The GPU is held
to the restatement in tests/test_gpu_nesting.py."""
import math
import types

import numpy as np
import pytest

import oracle as O
from dgcom_amd import atlas as A, cases, multirate as M, nesting as NS

LANDING_SLACK = 1e-9
KNOWN = {"cubic": (10845, 2285), "linear": (10862, 2293)}


def _global_steps(orc, x, c, t_end):
    t, n = 0.0, 0
    while t < t_end:
        dt = orc.dt(*x, c.cfl, c.scheme)
        landed = t + dt * (1.0 + LANDING_SLACK) >= t_end
        if landed:
            dt = t_end - t
        orc.run_dts(*x, t, [dt], c.scheme)
        t = t_end if landed else t + dt
        n += 1
    return n


def test_the_gate_domain():
    c, pm = NS.gate_case("cubic")
    assert c.n_elements == 640 and int(pm.band_elements.sum()) == 48
    assert sum(pm.on_boundary) == 24 and (pm.nesting.slot != 0xFFFFFFFF).sum() == 408
    assert pm.snapped == 0
    assert 32.0 < pm.nesting.depth.min() and pm.nesting.depth.max() < 42.0
    # the bed blended towards the parent across the band, unchanged beyond it
    x, _ = cases.node_coordinates(c.mesh, c.order)
    plain = -(30.0 - 32.0 * x.ravel() / 40_000.0)
    far = x.ravel() > 3000.0
    assert np.array_equal(c.bed[far], plain[far])
    assert np.all(c.bed[x.ravel() == 0.0] < plain[x.ravel() == 0.0])


@pytest.mark.parametrize("variant", ["cubic", "linear"])
def test_the_nesting_gate_known_answers(variant):
    c, pm = NS.gate_case(variant)
    orc = O.Oracle.from_case(c)
    orc.set_nesting(pm.nesting)
    orc.set_nesting_boundary(True)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    assert _global_steps(orc, x, c, NS.GATE_HOURS * 3600.0) == KNOWN[variant][0]
    assert orc.clips() == 0


@pytest.mark.parametrize("variant", ["cubic", "linear"])
def test_the_nesting_gate_with_local_time_stepping(variant):
    c, pm = NS.gate_case(variant)
    orc = O.Oracle.from_case(c)
    orc.set_nesting(pm.nesting)
    orc.set_nesting_boundary(True)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    _, n, _ = M.run(orc, M.neighbours(c.mesh.links), x, 0.0, NS.GATE_HOURS * 3600.0, c.cfl, 3, c.scheme)
    assert n == KNOWN[variant][1]


def test_band_distances_are_to_the_tagged_faces():
    """A 4 km x 1 km channel open at the west: a node's distance is its x within the band."""
    mesh = cases.structured_mesh(0.0, 4000.0, 0.0, 1000.0, 4, 1, sides=(cases.WALL, cases.WALL, cases.WALL, cases.TIDE))
    case = types.SimpleNamespace(mesh=mesh, order=2, n_elements=mesh.n_elements, nn=9)
    d = NS.boundary_distances(case, 1500.0)
    x, _ = cases.node_coordinates(mesh, 2)
    for xx, dd in zip(x.ravel(), d):
        assert abs(dd - xx) < 1e-9 if xx <= 1500.0 else dd == math.inf
    assert NS.point_segment_distance((0.0, 1.0), (-1.0, 0.0), (1.0, 0.0)) == 1.0
    assert NS.point_segment_distance((3.0, 4.0), (-1.0, 0.0), (0.0, 0.0)) == 5.0


def test_stencils_distances_and_the_east_axis():
    grid = NS.GeoGrid([8.0, 8.1, 8.2], [63.0, 63.1])
    wet = [True, True, False, True, False, True]
    idx, w = NS.bilinear_stencil(grid, grid.locate(8.05, 63.05), lambda k: wet[k])
    # corners (0, 1, 3, 4): the dry one keeps weight 0 and points at a wet corner
    assert idx == [0, 1, 3, 0] and abs(sum(w) - 1.0) < 1e-15 and w[3] == 0.0
    assert abs(w[0] - 1.0 / 3.0) < 1e-12
    assert NS.bilinear_stencil(grid, grid.locate(8.15, 63.05), lambda k: False) is None
    assert grid.locate(8.25, 63.05) is None
    assert abs(NS.geodesic_distance((8.0, 63.0), (8.0, 63.1)) - 11_119.5) < 1.0
    c, s = NS.east_axis(A.LocalProjection(63.0, 8.0), 63.0, 8.0)
    assert abs(c - 1.0) < 1e-12 and abs(s) < 1e-12
