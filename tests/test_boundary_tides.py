"""Tides per boundary node (BoundaryTides, src/boundary/tidal_atlas.rs): a
uniform harmonic tide written as per-node constants (C, S at the epoch's nodal
terms, omega' with the nodal phase drift) gives the harmonic tide's boundary
fluxes to rounding over a 25 h run of nodal drift; with a velocity, the boundary
mass flux is the characteristic state's closed form (Riemann invariants from the
interior at rest and from the external state); the transport scaling follows
u D_source / D_child, limited. The GPU build is held to the restatement in
tests/test_gpu_boundary_tides.py."""
import math

import numpy as np
import pytest

import oracle as O
from dgcom_amd import cases

G = 9.81


def test_the_uniform_tide_as_node_constants():
    c = cases.tidal_channel()
    orc_h = O.Oracle.from_case(c)
    orc_a = O.Oracle.from_case(c)
    orc_a.set_boundary_tides(cases.boundary_tides_of(c, c.tide))
    rng = np.random.default_rng(1)
    h = c.h + 0.1 * (c.h > 0.5) * rng.standard_normal(c.h.size)
    hu = 0.2 * rng.standard_normal(c.h.size) * (h > 0.5)
    hv = 0.2 * rng.standard_normal(c.h.size) * (h > 0.5)
    h = np.maximum(h, 0.0)
    nodes = cases.open_face_nodes(c)
    n1 = c.order + 1
    idx = [(k * 4 + f) * n1 + fi for k, f, fi, _, _ in nodes]
    for t in (0.0, 1800.0, 7300.0, 45_000.0, 90_000.0):
        a = orc_h.face_mass(h, hu, hv, t).ravel()[idx]
        b = orc_a.face_mass(h, hu, hv, t).ravel()[idx]
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(a).max()), (t, np.abs(a - b).max())
        ra, rb = orc_h.rhs(h, hu, hv, t), orc_a.rhs(h, hu, hv, t)
        for x, y in zip(ra, rb):
            assert np.abs(x - y).max() <= 1e-10 * max(1.0, np.abs(x).max())


def _channel(depth=5.0):
    mesh = cases.structured_mesh(0.0, 400.0, 0.0, 100.0, 4, 1, sides=(cases.WALL, cases.WALL, cases.WALL, cases.TIDE))
    x, _ = cases.node_coordinates(mesh, 2)
    bed = np.full(x.size, -depth)
    h, hu, hv = cases.still_water(bed)
    return cases.Case("channel", 2, mesh, bed, np.zeros(mesh.n_elements), 0.0, cases.Tide(), h, hu, hv)


@pytest.mark.parametrize("u_in", [0.5, -0.4, 9.0])
def test_an_external_velocity_through_the_characteristic_boundary(u_in):
    # west boundary (outward normal -x), interior at rest, surface 0 outside:
    # w_out = 2 c, w_in = -u_in - 2 c, so subcritically c_b = c + u_in / 4 and
    # u_n,b = -u_in / 2; supercritical inflow (u_in > c) takes the external state
    c = _channel()
    bt = cases.boundary_tides(c, [1e-4], [0.0], lambda x, y: np.array([[0.0, 0.0, u_in, 0.0, 0.0, 0.0]]),
                              velocity=True)
    orc = O.Oracle.from_case(c, friction=False)
    orc.set_boundary_tides(bt)
    fm = orc.face_mass(c.h, c.hu, c.hv, 0.0).ravel()
    n1 = 3
    west = [(k * 4 + 3) * n1 + fi for k in range(c.n_elements) for fi in range(n1)
            if c.mesh.links[k, 3] == cases.BOUNDARY]
    ci = math.sqrt(G * 5.0)
    if u_in > ci:
        h_b, un_b = 5.0, -u_in
    else:
        cb = ci + u_in / 4.0
        h_b, un_b = cb * cb / G, -u_in / 2.0
    assert np.allclose(fm[west], h_b * un_b, rtol=1e-13, atol=1e-13), (fm[west], h_b * un_b)


def test_the_transport_scaling():
    # u D_source / D_child limited to [1/max, max]: a source 12 m deep over a
    # 5 m bed scales the current 2.4x, clamped at 2
    c = _channel()
    u_in = 0.3
    out = []
    for ratio in (0.0, 2.0, 3.0):
        bt = cases.boundary_tides(c, [1e-4], [0.0], lambda x, y: np.array([[0.0, 0.0, u_in, 0.0, 0.0, 0.0]]),
                                  velocity=True, max_ratio=ratio, source_depth=lambda x, y: 12.0)
        orc = O.Oracle.from_case(c, friction=False)
        orc.set_boundary_tides(bt)
        fm = orc.face_mass(c.h, c.hu, c.hv, 0.0).ravel()
        k0 = next(k for k in range(c.n_elements) if c.mesh.links[k, 3] == cases.BOUNDARY)
        out.append(fm[(k0 * 4 + 3) * 3])
    ci = math.sqrt(G * 5.0)
    expect = [(ci + s * u_in / 4.0) ** 2 / G * (-s * u_in / 2.0) for s in (1.0, 2.0, 12.0 / 5.0)]
    assert np.allclose(out, expect, rtol=1e-13)
