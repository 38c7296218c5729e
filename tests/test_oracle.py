"""Pinning the CPU restatement (the oracle) of the reference's 2D shallow-water path.

Each check is one that a plausible mistake (a dropped term, a sign, a
transposed operand) fails; the planted-mistake builds at the end show it.

- closed forms: GLL points and weights for N = 1..3; the differentiation matrix
  exact on polynomials of degree <= N and summation-by-parts;
- the reference's published known answer: the manufactured nonlinear solution
  over a bed (docs/accuracy.md:175-185, tests/convergence_test.rs:961-1162),
  L2(h) at 4/8/16 elements per side, P2 and P3, to the three digits published;
- invariants with the reference's own bars (swe_2d_split_form.rs:1259-1477,
  docs/accuracy.md:201-211): lake at rest across shorelines, on rough beds and on
  jittered meshes; mass conservation; exactly opposite face mass fluxes; free
  stream on distorted periodic meshes;
- special cases: the Manning cube root within 2 ulp of the library cube root
  (friction.rs:631-982), the tide ramp (harmonic_tide.rs:158-165), positivity
  with no clips under the time-step bound.
"""
import math

import numpy as np
import pytest

from dgcom_amd import cases
import oracle as O


# --- closed forms ---------------------------------------------------------------

@pytest.mark.parametrize("order", [1, 2, 3])
def test_gll_points_and_weights(order):
    x, w, _ = O.operators(order)
    ref_x = {1: [-1, 1], 2: [-1, 0, 1], 3: [-1, -math.sqrt(0.2), math.sqrt(0.2), 1]}[order]
    ref_w = {1: [1, 1], 2: [1 / 3, 4 / 3, 1 / 3], 3: [1 / 6, 5 / 6, 5 / 6, 1 / 6]}[order]
    assert np.allclose(x, ref_x, rtol=0, atol=1e-15)
    assert np.allclose(w, ref_w, rtol=0, atol=1e-15)
    assert abs(w.sum() - 2.0) < 1e-15


@pytest.mark.parametrize("order", [1, 2, 3])
def test_differentiation_matrix(order):
    x, w, d = O.operators(order)
    for k in range(order + 1):
        exact = k * x ** (k - 1) if k else np.zeros_like(x)
        assert np.abs(d @ x ** k - exact).max() < 1e-13
    q = np.diag(w) @ d
    b = np.zeros_like(q)
    b[0, 0], b[-1, -1] = -1.0, 1.0
    assert np.abs(q + q.T - b).max() < 1e-14


# --- the reference's known answer -----------------------------------------------

def manufactured_error(order, n, cfl):
    """L2 depth error of the split form with entropy-stable faces on an n x n periodic
    mesh of [0, sqrt 2]^2 after t = 0.1 (tests/convergence_test.rs:1041-1097)."""
    length = math.sqrt(2.0)
    mesh = cases.structured_mesh(0.0, length, 0.0, length, n, n, periodic_x=True, periodic_y=True)
    x, y = cases.node_coordinates(mesh, order)
    s2 = math.sqrt(2.0)
    ox = 2.0 * math.pi * s2

    def bed(x, y):
        return 2.0 + 0.5 * np.sin(s2 * math.pi * x) + 0.5 * np.sin(s2 * math.pi * y)

    def depth(x, y, t):
        return 7.0 + np.cos(ox * x) * np.sin(ox * y) * math.cos(2.0 * math.pi * t) - bed(x, y)

    h = depth(x, y, 0.0).ravel()
    q = [h, 0.5 * h, 1.5 * h]
    orc = O.Oracle(order, mesh, bed(x, y).ravel(), surface=O.SURFACE_ES, subcells=False, friction=False,
                   wet_dry=False, manufactured=True)
    geo = orc.geometry()
    det = geo["det_j"]
    rx, ry, sx, sy = (geo[k] / det for k in ("jrx", "jry", "jsx", "jsy"))
    c = np.sqrt(9.81 * h)
    u, v = q[1] / h, q[2] / h
    rate = 0.25 * ((np.abs(u * rx + v * ry) + c * np.hypot(rx, ry)) + (np.abs(u * sx + v * sy) + c * np.hypot(sx, sy)))
    n_steps = math.ceil(0.1 / (cfl / ((2 * order + 1) * rate.max())))
    dt = 0.1 / n_steps
    t = 0.0
    for _ in range(n_steps):  # SSP-RK3 as in the reference test (axpy/scale form)
        q0 = [a.copy() for a in q]
        r = orc.rhs(*q, t)
        q = [a + dt * b for a, b in zip(q, r)]
        r = orc.rhs(*q, t + dt)
        q = [a + dt * b for a, b in zip(q, r)]
        q = [0.25 * a for a in q]
        q = [a + 0.75 * b for a, b in zip(q, q0)]
        r = orc.rhs(*q, t + 0.5 * dt)
        q = [a + dt * b for a, b in zip(q, r)]
        q = [(2.0 / 3.0) * a for a in q]
        q = [a + (1.0 / 3.0) * b for a, b in zip(q, q0)]
        t += dt
    _, w, _ = O.operators(order)
    w2 = np.outer(w, w).ravel()
    diff = q[0].reshape(-1, (order + 1) ** 2) - depth(x, y, 0.1)
    return math.sqrt(np.sum(w2 * diff * diff * det.reshape(diff.shape)))


def sig3(v):
    return float(f"{v:.2e}")


@pytest.mark.parametrize("order,cfl,published", [
    (2, 0.1, {4: 2.64e-1, 8: 3.94e-2, 16: 4.74e-3}),
    (3, 0.05, {4: 3.98e-2, 8: 3.17e-3, 16: 1.98e-4}),
])
def test_manufactured_solution_matches_published_table(order, cfl, published):
    errors = {n: manufactured_error(order, n, cfl) for n in published}
    for n, v in published.items():
        assert sig3(errors[n]) == v, (n, errors[n], v)
    observed = math.log2(errors[8] / errors[16])
    assert observed > order + 0.5


# --- invariants -----------------------------------------------------------------

def lake(order, jitter, rough=False, seed=3):
    c = cases.drying_beach(order=order, jitter=jitter, seed=seed)
    bed = c.bed.copy()
    if rough:  # face-discontinuous random bed of +-15 m around the slope, partly above the lake
        rng = np.random.default_rng(seed)
        bed = bed + rng.uniform(-1.5, 1.5, bed.size)
    h, hu, hv = cases.still_water(bed)
    return c, bed, h, hu, hv


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("jitter", [0.0, 0.2])
@pytest.mark.parametrize("rough", [False, True])
def test_lake_at_rest_across_shorelines(order, jitter, rough):
    c, bed, h, hu, hv = lake(order, jitter, rough)
    orc = O.Oracle(order, c.mesh, bed, manning=c.manning, f0=c.f0)  # open boundary at the still level
    r = orc.rhs(h, hu, hv, 0.0)
    nd, es = orc.masks(h)
    assert nd.sum() > 0 and es.sum() > 0  # shorelines and subcell elements present
    assert np.abs(r[0]).max() < 1e-12
    assert max(np.abs(r[1]).max(), np.abs(r[2]).max()) < 1e-9
    # and over steps: the surface stays at rest
    te = orc.run(h, hu, hv, 0.0, 20, 0.5)
    eta = h + bed
    wet = h > 1e-3
    assert te > 0 and orc.clips() == 0
    assert np.abs(eta[wet]).max() < 1e-13
    assert np.abs(hu).max() < 1e-12 and np.abs(hv).max() < 1e-12


def walled_basin(order, jitter, seed=11, flow=0.2):
    c = cases.drying_beach(order=order, jitter=jitter, seed=seed)
    mesh = c.mesh
    mesh.tags[:] = cases.WALL
    rng = np.random.default_rng(seed)
    h, hu, hv = cases.still_water(c.bed)
    h = np.where(h > 0.0, h + 0.05 * rng.standard_normal(h.size), 0.0).clip(min=0.0)
    hu = flow * rng.standard_normal(h.size) * h
    hv = flow * rng.standard_normal(h.size) * h
    return c, h, hu, hv


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_mass_conservation_and_opposite_face_fluxes(order, jitter):
    c, h, hu, hv = walled_basin(order, jitter)
    orc = O.Oracle(order, c.mesh, c.bed, manning=c.manning, f0=c.f0)
    geo = orc.geometry()
    r = orc.rhs(h, hu, hv, 0.0)
    mass = np.sum(geo["mass"] * h)
    assert abs(np.sum(geo["mass"] * r[0])) / mass < 1e-14
    fm = orc.face_mass(h, hu, hv, 0.0)
    links = c.mesh.links
    n1 = order + 1
    for k in range(c.n_elements):
        for f in range(4):
            l = int(links[k, f])
            if l == cases.BOUNDARY:
                continue
            nb, nf = l >> 4, l & 3
            assert np.array_equal(fm[k, f], -fm[nb, nf][::-1])
    # and over steps: mass to round-off, no clips
    m0 = np.sum(geo["mass"] * h)
    orc.run(h, hu, hv, 0.0, 30, 0.5)
    assert abs(np.sum(geo["mass"] * h) - m0) / m0 < 1e-13
    assert orc.clips() == 0 and h.min() >= 0.0


@pytest.mark.parametrize("order", [1, 2, 3])
def test_free_stream_on_distorted_periodic_mesh(order):
    n, length = 6, 600.0
    tau = 2 * math.pi / length
    a = 0.04 * length

    def distort(x, y):
        sx, sy = np.sin(tau * x), np.sin(tau * y)
        return x + a * sx * sy + 0.5 * a * np.sin(tau * y), y - 0.7 * a * sx * sy + 0.4 * a * np.sin(tau * x)

    mesh = cases.structured_mesh(0, length, 0, length, n, n, periodic_x=True, periodic_y=True, distort=distort)
    nn = (order + 1) ** 2
    h = np.full(mesh.n_elements * nn, 10.0)
    hu, hv = 0.3 * h, -0.2 * h
    orc = O.Oracle(order, mesh, np.full(h.size, -10.0), friction=False)
    r = orc.rhs(h, hu, hv, 0.0)
    scale = 9.81 * 10.0 ** 2 / (length / n)
    assert max(np.abs(x).max() for x in r) < 1e-12 * scale


# --- special cases --------------------------------------------------------------

def test_manning_cube_root_within_two_ulp():
    rng = np.random.default_rng(0)
    xs = np.concatenate([10.0 ** rng.uniform(-6, 4, 20000), [1e-6, 1.0, 8.0, 27.0, 1000.0]])
    for x in xs:
        got = O.cbrt(float(x))
        ref = np.cbrt(x)
        assert abs(got - ref) <= 2 * np.spacing(ref)


def test_tide_ramp_and_elevation():
    mesh = cases.structured_mesh(0, 1, 0, 1, 1, 1)
    orc = O.Oracle(1, mesh, np.zeros(4))
    orc.set_tide(cases.Tide(amplitude=(1.0,), omega=(0.0,), phase=(0.0,), ramp=10.0))
    assert orc.tide_elevation(-1.0) == 0.0
    assert orc.tide_elevation(5.0) == 0.5
    assert orc.tide_elevation(20.0) == 1.0
    # phase is minus the Greenwich lag: 0.5 m M2 with a 90 degree lag peaks a quarter period later
    period = cases.M2_PERIOD_S
    orc.set_tide(cases.Tide(amplitude=(0.5,), omega=(2 * math.pi / period,), phase=(-math.pi / 2,), mean=0.1, ramp=0.0))
    assert abs(orc.tide_elevation(0.0) - 0.1) < 1e-12
    assert abs(orc.tide_elevation(0.25 * period) - 0.6) < 1e-12


def test_tidal_estuary_runs_without_clips():
    c = cases.tidal_estuary(nx=20, ny=4)
    orc = O.Oracle.from_case(c)
    h, hu, hv = c.h.copy(), c.hu.copy(), c.hv.copy()
    te = orc.run(h, hu, hv, 0.0, 200, c.cfl)
    assert te > 0 and np.isfinite(h).all() and orc.clips() == 0


# --- planted mistakes: each build must fail the check that pins its term ---------

@pytest.mark.parametrize("mutant,check", [
    ("NO_BED", "lake"),
    ("TRANSPOSE", "lake"),
    ("LIFT", "mass"),
])
def test_planted_mistakes_are_caught(mutant, check):
    L = O.mutant(mutant)
    with O.using(L):
        if check == "lake":
            c, bed, h, hu, hv = lake(2, 0.2)
            orc = O.Oracle(2, c.mesh, bed, manning=c.manning, f0=c.f0, tide=c.tide)
            r = orc.rhs(h, hu, hv, 0.0)
            assert max(np.abs(r[1]).max(), np.abs(r[2]).max()) > 1e-6
        else:
            c, h, hu, hv = walled_basin(2, 0.2)
            orc = O.Oracle(2, c.mesh, c.bed, manning=c.manning, f0=c.f0)
            geo = orc.geometry()
            r = orc.rhs(h, hu, hv, 0.0)
            assert abs(np.sum(geo["mass"] * r[0])) / np.sum(geo["mass"] * h) > 1e-10
        del orc
