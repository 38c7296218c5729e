"""The inversion ensemble (uneven meshes, N in {2, 3}, each member with its own
truth/prior bed pair and 12 stations, wet/dry active): every member's misfit
gradient, taken in one sequence with the others', against central differences
of its own forward runs along random directions in (bed, Manning n), with
identical wet/dry masks at both perturbed ends; and each member's gradient
equals the one it gets alone, bit for bit.

Small members (a few hundred elements, 30 steps) keep the finite differences
quick; the directions are smooth bed fields (a per-node random bed direction
is not differentiable where two face traces tie) plus random Manning
perturbations. The members start from a jittered level and velocity: from rest
with continuous depths the HLL wave speeds tie at every face, the misfit has a
kink at the prior bed, and its central differences are not its gradient."""
import numpy as np
import pytest

from dgcom_amd import cases, inversion, swe2d

N_DIRECTIONS = 8
SMALL = dict(n_members=3, k_lo=300, k_hi=900, orders=(2, 3), n_stations=12, every=5, n_steps=30, seed=31,
             correlation=800.0, mean_depth=20.0)


@pytest.fixture(scope="module")
def ensemble():
    members = inversion.ensemble_workload(**SMALL)
    physics = []
    for inv in members:
        phys = swe2d.SWEPhysics2D.from_case(inv.case)
        inversion.observe(inv, phys)
        physics.append(phys)
    grads = inversion.ensemble_gradients(members, physics, snapshots=4)
    return members, physics, grads


def _misfit(inv, bed, manning):
    c = inv.case
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_bed(bed)
    phys.set_manning(manning)
    phys.set_stations(inv.stations.elements, inv.stations.weights)
    tape = phys.tape(len(inv.dts) + 1)
    J = tape.forward(phys.state(c.h, c.hu, c.hv), 0.0, inv.dts, inv.every, inv.obs)
    tape.close()
    s = phys.state(c.h, c.hu, c.hv)
    t = 0.0
    for dt in inv.dts:
        phys.step(s, t, dt)
        t += dt
    return J, phys.masks(s)


def test_members_are_uneven_and_wet_dry():
    members = inversion.ensemble_workload(**SMALL)
    assert len({m.case.n_elements for m in members}) == 3
    assert {m.case.order for m in members} <= {2, 3}
    for m in members:
        assert m.stations.elements.size == 12
        assert (m.prior_bed > 0.0).mean() >= 0.01, "the member has no land: wet/dry would not be exercised"


@pytest.mark.gpu
def test_each_gradient_matches_central_differences(ensemble):
    members, physics, grads = ensemble
    for m, (inv, (J, gbed, gman)) in enumerate(zip(members, grads)):
        c = inv.case
        assert J > 0.0 and np.abs(gbed).max() > 0.0 and np.abs(gman).max() > 0.0
        x, y = cases.node_coordinates(c.mesh, c.order)
        L = float(c.mesh.corners[:, :, 0].max())
        rng = np.random.default_rng(100 + m)
        for d in range(N_DIRECTIONS):
            dbed = cases.gaussian_random_field(x, y, (0.0, L, 0.0, L), 1.0, 0.3 * L, 500 + 10 * m + d).ravel()
            dman = 0.002 * rng.standard_normal(c.n_elements)
            g = float(gbed @ dbed) + float(gman @ dman)
            eps = 1e-5
            Jp, mp = _misfit(inv, inv.prior_bed + eps * dbed, c.manning + eps * dman)
            Jm, mm = _misfit(inv, inv.prior_bed - eps * dbed, c.manning - eps * dman)
            assert np.array_equal(mp[0], mm[0]) and np.array_equal(mp[1], mm[1]), (m, d, "a wet/dry mask flipped")
            fd = (Jp - Jm) / (2.0 * eps)
            assert abs(fd - g) <= 1e-6 * abs(g), (m, d, fd, g)


@pytest.mark.gpu
def test_each_gradient_equals_its_own_run(ensemble):
    members, physics, grads = ensemble
    for inv, (J, gbed, gman) in zip(members, grads):
        c = inv.case
        phys = swe2d.SWEPhysics2D.from_case(c)
        phys.set_bed(inv.prior_bed)
        phys.set_stations(inv.stations.elements, inv.stations.weights)
        tape = phys.tape(4)
        J1 = tape.forward(phys.state(c.h, c.hu, c.hv), 0.0, inv.dts, inv.every, inv.obs)
        gb1, gm1 = tape.backward()
        tape.close()
        assert J1 == J and np.array_equal(gb1, gbed) and np.array_equal(gm1, gman)


@pytest.mark.gpu
def test_step_records_give_the_same_member_gradients(ensemble):
    """The members' tapes with step records (no reverse step recomputes its stages):
    every member's misfit and gradients bit for bit those of its tape without."""
    members, physics, grads = ensemble
    again = inversion.ensemble_gradients(members, physics, snapshots=4, keep_stages=True)
    for (J, gb, gm), (Jr, gbr, gmr) in zip(grads, again):
        assert J == Jr and np.array_equal(gb, gbr) and np.array_equal(gm, gmr)
