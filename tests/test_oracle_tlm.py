"""The oracle's tangent-linear model (its solver instantiated over dual numbers)
against central finite differences of the oracle itself.

The tangent model is what the adjoint's dot-product tests check against, so it
is pinned here first: along state directions, per-element Manning directions and
continuous bed directions (a bed that stays continuous across faces; the face
bed max(B-, B+) has a kink wherever the two traces are equal, so a per-node
random bed direction has no derivative there), for one stage and for a short run
over frozen time steps. Central differences at eps = 1e-4 carry an O(eps^2)
truncation error and are expected within 1e-7 relative.
"""
import numpy as np
import pytest

from dgcom_amd import cases
import oracle as O
from cases_for_tests import perturbed


def _case_and_state():
    c = cases.tidal_estuary(nx=6, ny=3)
    return c, perturbed(c, seed=1)


def _smooth_direction(c, seed):
    x, y = cases.node_coordinates(c.mesh, c.order)
    return cases.gaussian_random_field(x, y, (0.0, 10000.0, 0.0, 2000.0), 1.0, 800.0, seed).ravel()


def _rel(a, b):
    return max(np.abs(np.asarray(x) - np.asarray(y)).max() / max(np.abs(np.asarray(y)).max(), 1e-300)
               for x, y in zip(a, b))


@pytest.mark.parametrize("which", ["state", "manning", "bed"])
def test_stage_tangent_matches_finite_differences(which):
    c, base = _case_and_state()
    orc = O.Oracle.from_case(c)
    dt = orc.dt(*base, c.cfl)
    rng = np.random.default_rng(0)
    n = base[0].size
    dx = [rng.standard_normal(n) * 1e-3 for _ in range(3)] if which == "state" else [np.zeros(n)] * 3
    dbed = _smooth_direction(c, 7) if which == "bed" else np.zeros(n)
    dman = rng.standard_normal(c.n_elements) * 0.01 if which == "manning" else np.zeros(c.n_elements)
    tlm = O.TangentLinear.from_case(c)
    tlm.set_param_tangents(dbed, dman)
    # the second SSP-RK3 stage (a=3/4, b=1/4): base and input differ
    inp = orc.stage(base, base, 1.0, 0.0, dt, 0.0, False)
    _, dinp = tlm.stage(base, dx, base, dx, 1.0, 0.0, dt, 0.0, False)
    _, dout = tlm.stage(base, dx, inp, dinp, 0.75, 0.25, 0.25 * dt, 0.5 * dt, False)

    def f(s):
        o = O.Oracle(c.order, c.mesh, c.bed + s * dbed, manning=c.manning + s * dman, f0=c.f0, tide=c.tide)
        b = tuple(a + s * d for a, d in zip(base, dx))
        i = o.stage(b, b, 1.0, 0.0, dt, 0.0, False)
        return o.stage(b, i, 0.75, 0.25, 0.25 * dt, 0.5 * dt, False)

    eps = 1e-4
    fd = [(p - m) / (2 * eps) for p, m in zip(f(eps), f(-eps))]
    assert max(np.abs(d).max() for d in dout) > 0.0
    assert _rel(dout, fd) < 1e-7


def test_run_tangent_matches_finite_differences():
    c, base = _case_and_state()
    orc = O.Oracle.from_case(c)
    h, hu, hv = (a.copy() for a in base)
    _, dts = orc.run(h, hu, hv, 0.0, 5, c.cfl, record_dt=True)
    dts = np.asarray(dts)
    rng = np.random.default_rng(2)
    n = base[0].size
    dx = [rng.standard_normal(n) * 1e-3 for _ in range(3)]
    dbed = _smooth_direction(c, 9)
    dman = rng.standard_normal(c.n_elements) * 0.01
    tlm = O.TangentLinear.from_case(c)
    tlm.set_param_tangents(dbed, dman)
    x = [a.copy() for a in base]
    d = [a.copy() for a in dx]
    tlm.run(x, d, 0.0, dts)

    def f(s):
        o = O.Oracle(c.order, c.mesh, c.bed + s * dbed, manning=c.manning + s * dman, f0=c.f0, tide=c.tide)
        y = [a + s * b for a, b in zip(base, dx)]
        o.run_dts(*y, 0.0, dts)
        return y

    eps = 1e-4
    fd = [(p - m) / (2 * eps) for p, m in zip(f(eps), f(-eps))]
    assert _rel(d, fd) < 1e-7
