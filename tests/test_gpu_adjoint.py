"""The discrete adjoint on the GPU, through the C ABI.

- Parity: the stage VJP of the HIP library equals the same kernels run on host
  threads, bit for bit (both are built without contraction).
- Dot-product tests against the oracle's tangent-linear model (dual numbers):
  <w, J d> = <J^T w, d> for a stage, the RHS and a whole SSP-RK3 step, to 1e-12.
- The station misfit's gradient with respect to the nodal bed and per-element
  Manning n, through binomial checkpointing: identical bits for every snapshot
  count, the tangent-linear directional derivative to 1e-10, and central finite
  differences of the GPU forward to 1e-6 along smooth bed and Manning
  directions, with identical wet/dry masks at both perturbed ends.
"""
import numpy as np
import pytest

from dgcom_amd import cases, stations, swe2d
import oracle as O
from cases_for_tests import perturbed, small_unstructured
from emu_bridge import EmulatedKernels
from test_adjoint_emulated import STAGE_CASES

pytestmark = pytest.mark.gpu


def _dot(a, b):
    return sum(float(np.dot(x, y)) for x, y in zip(a, b))


def _stage_inputs(c, seed=1):
    orc = O.Oracle.from_case(c)
    base = perturbed(c, seed=seed)
    dt = orc.dt(*base, c.cfl)
    inp = orc.stage(base, base, 1.0, 0.0, dt, 0.0, False)
    return base, inp, dt


def _gpu_stage_vjp(phys, base, inp, w, a, b, cc, time, rhs_only=False):
    B, I, W = phys.state(*base), phys.state(*inp), phys.state(*w)
    gb, gi = phys.state().zero(), phys.state().zero()
    phys.param_grads_zero()
    if rhs_only:
        phys.rhs_vjp(I, time, W, gi)
    else:
        phys.stage_vjp(B, I, a, b, cc, time, W, gb, gi)
    gbed, gman = phys.param_grads()
    return gb.download(), gi.download(), gbed, gman


def _estuary_with_unmatched_normals():
    # an interior element's vertex moved for it alone: the edges at that vertex have two
    # normals, and the library runs them through its split-edge kernel
    c = cases.tidal_estuary(nx=6, ny=3)
    c.mesh.corners[7, 2] += (3.0, -2.0)
    return c


@pytest.mark.parametrize("name,make", STAGE_CASES + [("estuary_unmatched_normals", _estuary_with_unmatched_normals)])
def test_stage_vjp_matches_emulated_kernels_bitwise(name, make):
    c = make()
    base, inp, dt = _stage_inputs(c)
    rng = np.random.default_rng(3)
    w = [rng.standard_normal(base[0].size) for _ in range(3)]
    phys = swe2d.SWEPhysics2D.from_case(c)
    got = _gpu_stage_vjp(phys, base, inp, w, 0.75, 0.25, 0.25 * dt, 0.5 * dt)
    want = EmulatedKernels(c).stage_vjp(base, inp, w, 0.75, 0.25, 0.25 * dt, 0.5 * dt)
    for g, e, what in zip(got, want, ("base", "input", "bed", "manning")):
        g, e = np.concatenate(g) if isinstance(g, list) else g, np.concatenate(e) if isinstance(e, list) else e
        assert np.array_equal(g, e), (what, np.abs(g - e).max())


@pytest.mark.parametrize("name,make", STAGE_CASES)
@pytest.mark.parametrize("mode", ["rhs", "stage"])
def test_stage_vjp_dot_product_against_tangent_linear(name, make, mode):
    c = make()
    base, inp, dt = _stage_inputs(c)
    rng = np.random.default_rng(7)
    n = base[0].size
    w = [rng.standard_normal(n) for _ in range(3)]
    phys = swe2d.SWEPhysics2D.from_case(c)
    if mode == "rhs":
        a, b, cc, time, inp = 0.0, 0.0, 0.0, 0.0, base
    else:
        a, b, cc, time = 0.75, 0.25, 0.25 * dt, 0.5 * dt
    gb, gi, gbed, gman = _gpu_stage_vjp(phys, base, inp, w, a, b, cc, time, rhs_only=mode == "rhs")
    dxb = [rng.standard_normal(n) for _ in range(3)] if mode == "stage" else [np.zeros(n)] * 3
    dxi = [rng.standard_normal(n) for _ in range(3)]
    dbed = rng.standard_normal(n)
    dman = rng.standard_normal(c.n_elements)
    tlm = O.TangentLinear.from_case(c)
    tlm.set_param_tangents(dbed, dman)
    if mode == "rhs":
        _, dout = tlm.rhs(inp, dxi, time)
    else:
        _, dout = tlm.stage(base, dxb, inp, dxi, a, b, cc, time, False)
    lhs = _dot(w, dout)
    rhs = _dot(gb, dxb) + _dot(gi, dxi) + float(gbed @ dbed) + float(gman @ dman)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("name,make", [STAGE_CASES[0], STAGE_CASES[2], STAGE_CASES[4]])
def test_step_vjp_dot_product_against_tangent_linear(name, make):
    c = make()
    x = perturbed(c, seed=2)
    dt = O.Oracle.from_case(c).dt(*x, c.cfl)
    t = 100.0
    rng = np.random.default_rng(9)
    n = x[0].size
    w = [rng.standard_normal(n) for _ in range(3)]
    phys = swe2d.SWEPhysics2D.from_case(c)
    X, W, G = phys.state(*x), phys.state(*w), phys.state()
    phys.param_grads_zero()
    phys.step_vjp(X, t, dt, W, G)
    g = G.download()
    gbed, gman = phys.param_grads()
    dx = [rng.standard_normal(n) for _ in range(3)]
    dbed = rng.standard_normal(n)
    dman = rng.standard_normal(c.n_elements)
    tlm = O.TangentLinear.from_case(c)
    tlm.set_param_tangents(dbed, dman)
    xv = [a.copy() for a in x]
    dv = [a.copy() for a in dx]
    tlm.run(xv, dv, t, [dt])
    lhs = _dot(w, dv)
    rhs = _dot(g, dx) + float(gbed @ dbed) + float(gman @ dman)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


# --- the station misfit ------------------------------------------------------

N_STEPS, EVERY = 40, 5


class Problem:
    """A shallow jittered basin with shorelines (wet/dry active), 6 stations,
    observations from a truth bed, a prior bed that differs by a smooth field."""

    def __init__(self):
        c = small_unstructured(2)
        self.case = c
        x, y = cases.node_coordinates(c.mesh, c.order)
        ext = (0.0, 2400.0, 0.0, 2400.0)
        self.truth_bed = c.bed.copy()
        self.prior_bed = c.bed + 0.5 * cases.gaussian_random_field(x, y, ext, 1.0, 400.0, 21).ravel()
        self.smooth_dir = cases.gaussian_random_field(x, y, ext, 1.0, 300.0, 33).ravel()
        self.st = stations.random_stations(c, 6, seed=5, min_depth=2.0)
        self.x0 = perturbed(c, seed=4, amp=0.02)
        phys = swe2d.SWEPhysics2D.from_case(c)
        phys.set_bed(self.prior_bed)
        s = phys.state(*self.x0)
        _, self.dts = phys.run(s, 0.0, N_STEPS, c.cfl, record_dt=True)
        # observations: the truth bed run at the same steps
        phys.set_bed(self.truth_bed)
        s = phys.state(*self.x0)
        phys.set_stations(self.st.elements, self.st.weights)
        obs, t = [], 0.0
        for i, dt in enumerate(self.dts):
            phys.step(s, t, dt)
            t += dt
            if (i + 1) % EVERY == 0:
                obs.append(phys.sample_stations(s))
        self.obs = np.array(obs)

    def physics(self, bed=None, manning=None):
        phys = swe2d.SWEPhysics2D.from_case(self.case)
        phys.set_bed(self.prior_bed if bed is None else bed)
        if manning is not None:
            phys.set_manning(manning)
        phys.set_stations(self.st.elements, self.st.weights)
        return phys

    def misfit(self, bed=None, manning=None, snapshots=N_STEPS + 1):
        phys = self.physics(bed, manning)
        tape = phys.tape(snapshots)
        J = tape.forward(phys.state(*self.x0), 0.0, self.dts, EVERY, self.obs)
        nd, es = phys.masks(tape_final_state(phys, self, bed, manning))
        return J, (nd, es)

    def gradient(self, snapshots, keep_stages=False):
        phys = self.physics()
        tape = phys.tape(snapshots)
        if keep_stages:
            tape.keep_stages()
        J = tape.forward(phys.state(*self.x0), 0.0, self.dts, EVERY, self.obs)
        gx0 = phys.state()
        gbed, gman = tape.backward(1.0, gx0)
        if keep_stages:
            self.records_used = tape.records_used()
        return J, gbed, gman, gx0.download(), tape.stats()


def tape_final_state(phys, prob, bed, manning):
    s = phys.state(*prob.x0)
    t = 0.0
    for dt in prob.dts:
        phys.step(s, t, dt)
        t += dt
    return s


@pytest.fixture(scope="module")
def problem():
    return Problem()


def test_misfit_gradient_is_independent_of_the_checkpoint_schedule(problem):
    J0, gb0, gm0, gx0, (fwd0, vjp0) = problem.gradient(N_STEPS + 1)
    assert fwd0 == N_STEPS and vjp0 == N_STEPS
    for snaps in (5, 2, 1):
        J, gb, gm, gx, (fwd, vjp) = problem.gradient(snaps)
        assert J == J0
        assert np.array_equal(gb, gb0) and np.array_equal(gm, gm0) and all(np.array_equal(a, b) for a, b in zip(gx, gx0))
        assert vjp == N_STEPS
        assert fwd == swe2d.revolve_count(N_STEPS, snaps)


@pytest.mark.parametrize("snaps", [N_STEPS + 1, 5, 2, 1])
def test_step_records_give_the_same_gradient_without_recomputing(problem, snaps):
    """A tape keeping step records (every step's with a snapshot per step, else
    snapshots + 1 made by the sweep and the schedule's advances) reverses every step
    from its record: the same misfit and gradients bit for bit as the tape without,
    the same forward steps (the schedule's), no reverse step left to recompute."""
    J0, gb0, gm0, gx0, (fwd0, _) = problem.gradient(snaps)
    J, gb, gm, gx, (fwd, vjp) = problem.gradient(snaps, keep_stages=True)
    assert J == J0 and fwd == fwd0 == swe2d.revolve_count(N_STEPS, snaps) and vjp == N_STEPS
    assert problem.records_used == N_STEPS
    assert np.array_equal(gb, gb0) and np.array_equal(gm, gm0) and all(np.array_equal(a, b) for a, b in zip(gx, gx0))


def test_misfit_gradient_matches_tangent_linear(problem):
    c = problem.case
    J, gbed, gman, gx0, _ = problem.gradient(6)
    assert J > 0.0 and np.abs(gbed).max() > 0.0 and np.abs(gman).max() > 0.0
    rng = np.random.default_rng(13)
    n = c.h.size
    dbed = rng.standard_normal(n)
    dman = 0.002 * rng.standard_normal(c.n_elements)
    dx0 = [1e-3 * rng.standard_normal(n) for _ in range(3)]
    tlm = O.TangentLinear(c.order, c.mesh, problem.prior_bed, manning=c.manning, f0=c.f0, g=c.g, h_min=c.h_min,
                          h_dry=c.h_dry, tide=c.tide)
    tlm.set_param_tangents(dbed, dman)
    x = [a.copy() for a in problem.x0]
    dx = [a.copy() for a in dx0]
    dJ, t = 0.0, 0.0
    for m in range(N_STEPS // EVERY):
        seg = problem.dts[m * EVERY:(m + 1) * EVERY]
        t = tlm.run(x, dx, t, seg)
        eta = problem.st.eta(x[0], problem.prior_bed)
        deta = problem.st.eta(dx[0], dbed)
        dJ += float(np.sum(2.0 * (eta - problem.obs[m]) * deta))
    adj = float(gbed @ dbed) + float(gman @ dman) + _dot(gx0, dx0)
    assert abs(adj - dJ) <= 1e-10 * abs(dJ), (adj, dJ)


@pytest.mark.parametrize("param", ["bed", "manning"])
def test_misfit_gradient_matches_central_differences(problem, param):
    c = problem.case
    J, gbed, gman, _, _ = problem.gradient(6)
    if param == "bed":
        d = problem.smooth_dir
        g = float(gbed @ d)
        eps = 1e-5
        Jp, mp = problem.misfit(bed=problem.prior_bed + eps * d)
        Jm, mm = problem.misfit(bed=problem.prior_bed - eps * d)
    else:
        rng = np.random.default_rng(17)
        d = 0.002 * rng.standard_normal(c.n_elements)
        g = float(gman @ d)
        eps = 1e-3
        Jp, mp = problem.misfit(manning=c.manning + eps * d)
        Jm, mm = problem.misfit(manning=c.manning - eps * d)
    assert np.array_equal(mp[0], mm[0]) and np.array_equal(mp[1], mm[1]), "a wet/dry mask flipped"
    fd = (Jp - Jm) / (2.0 * eps)
    assert abs(fd - g) <= 1e-6 * abs(g), (fd, g)
