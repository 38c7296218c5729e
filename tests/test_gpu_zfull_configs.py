"""The configurations at their full sizes on the GPU.

- the idealised tidal estuary (100 x 20 affine quads, N = 2) over its whole 25 h of M2 tide, from rest,
  against the CPU restatement: every state bit and every time step equal;
- the batched forward ensemble: 64 members, K log-uniform in 20,000-400,000 jittered quads, N in {1, 2, 3},
  500 SSP-RK3 steps, graph-replayed and tuned: every member's state, time and wet/dry masks equal its own run's;
- the bathymetry inversion: 250,000 jittered quads, N = 2, 20 stations every 10 of 2000 steps, a
  20-checkpoint binomial tape: the misfit gradient against central differences along 32 seeded directions in
  (bed, Manning n), the projected gradient of each set of 8 to 1e-6 relative (1e-5 where a direction is smooth
  only at steps whose rounding noise exceeds that), with identical wet/dry masks at both perturbed ends;
- the inversion ensemble: 32 uneven members (20,000-250,000 quads, N in {2, 3}, wet/dry in each), 12 stations,
  1000 steps: each member's gradient against central differences along 8 directions, with identical end masks:
  the median error 1e-6, 90 % of them 1e-4, every one 1e-1 (some directions take other wet/dry branches during
  the runs at every practical step, and their differences straddle those kinks); at most 4 of the 256 without a
  step down to 1e-8 that keeps the end masks identical.

The directions are smooth bed fields (a per-node random bed direction is not differentiable where two face
traces tie) plus random Manning perturbations. A step of the differences is one at which both perturbed runs took
the same wet/dry branches at their ends (see _differences). The initial states carry the inversion
workload's jitter (from rest with depths continuous across faces the HLL wave speeds tie at every face and the
misfit has a kink).
"""
import numpy as np
import pytest

from dgcom_amd import cases, inversion, swe2d
import oracle as O

pytestmark = pytest.mark.gpu

# Steps of the central differences (bed in m, Manning in s m^-1/3, times the direction), tried in turn until
# the two perturbed runs end with the same wet/dry masks (see _differences). Their rounding noise
# (the two runs' different roundings over 1000-2000 steps) is divided by 2 eps, their truncation grows as eps^2.
EPS_INVERSION = (1e-5, 2.5e-6, 6.25e-7, 1.5625e-7, 3.90625e-8)
EPS_MEMBERS = (1e-5, 2.5e-6, 6.25e-7, 1.5625e-7, 3.90625e-8, 9.765625e-9)
REL = 1e-6          # the gradient's bar: the directions' projected gradient as a whole
# ... where a direction is smooth only below CLEAN_EPS: there the differences' rounding noise, measured at
# about 1e-13 / eps, exceeds 1e-6 of the derivative, and the set is held to REL_NOISY
CLEAN_EPS, REL_NOISY = 2.5e-6, 1e-5
REL_EACH = 1e-4     # each direction, against max(|g|, a tenth of the directions' mean |g|)
# The inversion ensemble's 256 directions: some runs take another wet/dry branch during the perturbed runs at
# every practical step (their differences straddle kinks and are not the derivative), so its bars are on the
# distribution of the directions' errors (against s as above): the median to REL, 90 % to REL_EACH, every one
# to REL_GROSS (a wrong sign or a dropped term).
REL_GROSS = 1e-1


def _misfit_and_masks(phys, inv, bed, manning, x0, every_sample=False):
    """J = sum over samples and stations of (eta - obs)^2 of a forward run at the recorded dts from x0, and the
    wet/dry masks at its end (every_sample: with a checksum of the masks at every sample before them)."""
    phys.set_bed(bed)
    phys.set_manning(manning)
    x0_h, x0_hu, x0_hv = x0
    s = phys.state(x0_h, x0_hu, x0_hv)
    rng = np.random.default_rng(1)
    wn, we = rng.standard_normal(phys.n_nodes_total), rng.standard_normal(phys.K)
    t, J, m, sums = 0.0, 0.0, 0, []
    for i, dt in enumerate(inv.dts):
        phys.step(s, t, dt)
        t += dt
        if (i + 1) % inv.every == 0:
            eta = phys.sample_stations(s)
            J += float(np.sum((eta - inv.obs[m]) ** 2))
            m += 1
            if every_sample:
                nd, es = phys.masks(s)
                sums.append((float(wn @ nd), float(we @ es)))
    nd, es = phys.masks(s)
    s.free()
    return J, (nd, es, sums)


def _direction(c, x, y, L, seed, d):
    dbed = cases.gaussian_random_field(x, y, (0.0, L, 0.0, L), 1.0, 0.3 * L, seed + 17 * d).ravel()
    dman = 0.002 * np.random.default_rng(seed + 17 * d).standard_normal(c.n_elements)
    return dbed, dman


def _central(phys, inv, dbed, dman, eps):
    c, x0 = inv.case, (inv.case.h, inv.case.hu, inv.case.hv)
    Jp, mp = _misfit_and_masks(phys, inv, inv.prior_bed + eps * dbed, c.manning + eps * dman, x0)
    Jm, mm = _misfit_and_masks(phys, inv, inv.prior_bed - eps * dbed, c.manning - eps * dman, x0)
    same_end = np.array_equal(mp[0], mm[0]) and np.array_equal(mp[1], mm[1])
    return (Jp - Jm) / (2.0 * eps), same_end


def _differences(inv, phys, gbed, gman, dirs, all_dirs, seed, tag, eps_ladder, allow_unmatched=False):
    """Central differences fd of the misfit along each direction d of `dirs`, at the first step of the ladder at
    which the two perturbed runs end with the same wet/dry masks, against the adjoint's g = <grad J, d>; with
    s = max(|g|, G / 10), G the mean |g| over `all_dirs` (a direction nearly orthogonal to the gradient has a
    small g and differences whose rounding noise is as large as every other direction's). Returns
    [(fd, g, eps, |fd - g| / s)]."""
    c = inv.case
    x, y = cases.node_coordinates(c.mesh, c.order)
    L = float(c.mesh.corners[:, :, 0].max())
    gs = {}
    for d in all_dirs:
        dbed, dman = _direction(c, x, y, L, seed, d)
        gs[d] = float(gbed @ dbed) + float(gman @ dman)
    G = float(np.mean([abs(v) for v in gs.values()]))
    out = []
    for d in dirs:
        dbed, dman = _direction(c, x, y, L, seed, d)
        g = gs[d]
        sc = max(abs(g), 0.1 * G)
        for eps in eps_ladder:
            fd, same_end = _central(phys, inv, dbed, dman, eps)
            if same_end:
                break
            print(f"{tag} direction {d}: eps {eps:g}: masks differ at the end", flush=True)
        if not same_end and allow_unmatched:
            out.append((fd, g, eps, float("nan")))
            print(f"{tag} direction {d}: unmatched (the end masks differ at every step)", flush=True)
            continue
        assert same_end, (tag, d, "the perturbed runs end with different wet/dry masks at every step")
        out.append((fd, g, eps, abs(fd - g) / sc))
        print(f"{tag} direction {d}: eps {eps:g} fd {fd:.9e} adjoint {g:.9e} relative {abs(fd - g) / abs(g):.2e} "
              f"(of s: {abs(fd - g) / sc:.2e})", flush=True)
    phys.set_bed(inv.prior_bed)
    phys.set_manning(c.manning)
    return out


def _check_set(res, tag):
    """||fd - g|| <= REL ||g|| over the set (REL_NOISY if a step is below CLEAN_EPS: there the differences'
    rounding noise, about 1e-13 / eps, exceeds 1e-6 of the derivative), and each |fd - g| <= REL_EACH s."""
    fd = np.array([r[0] for r in res])
    g = np.array([r[1] for r in res])
    steps = [r[2] for r in res]
    rel = float(np.linalg.norm(fd - g) / np.linalg.norm(g))
    clean = min(steps) >= CLEAN_EPS
    print(f"{tag}: ||fd - g|| / ||g|| = {rel:.2e} ({'clean' if clean else f'noise-limited: smallest step {min(steps):g}'})",
          flush=True)
    assert max(r[3] for r in res) <= REL_EACH, (tag, max(r[3] for r in res))
    assert rel <= (REL if clean else REL_NOISY), (tag, rel, min(steps))


def test_tidal_estuary_25_hours_equals_the_restatement():
    c = cases.tidal_estuary()
    assert (c.n_elements, c.order) == (2000, 2)
    phys = swe2d.SWEPhysics2D.from_case(c)
    # the steps that take the estuary from rest through 25 h (the first whose end is at or past 25 h)
    s = phys.state(c.h, c.hu, c.hv)
    probe = 100_000
    _, dts = phys.run(s, 0.0, probe, c.cfl, record_dt=True)
    n = int(np.searchsorted(np.cumsum(dts), 25 * 3600.0)) + 1
    assert 80_000 < n < probe, n
    s.upload(c.h, c.hu, c.hv)
    t_gpu = phys.run(s, 0.0, n, c.cfl)
    orc = O.Oracle.from_case(c)
    h, hu, hv = c.h.copy(), c.hu.copy(), c.hv.copy()
    t_cpu = orc.run(h, hu, hv, 0.0, n, c.cfl)
    assert t_gpu == t_cpu and t_gpu >= 25 * 3600.0, (t_gpu, t_cpu)
    got = s.download()
    for g, r in zip(got, (h, hu, hv)):
        assert np.array_equal(g, r), np.abs(g - r).max()
    assert orc.clips() == 0 and phys.negative_depth_clips() == 0
    # the tide did come in: at 25 h (2.01 M2 periods, near high water) the surface stands above 0.3 m
    # everywhere and the flow has reached the closed end
    eta = h + c.bed
    assert eta.min() > 0.3, eta.min()
    x, _ = cases.node_coordinates(c.mesh, c.order)
    assert np.abs(hu[x.ravel() > 9000.0]).max() > 1e-3


def test_ensemble_of_64_members_equals_their_own_runs():
    members = cases.ensemble_members(64, 99, 20_000, 400_000, 500)
    assert sorted({c.order for c in members}) == [1, 2, 3]
    n_steps = 500
    phys = [swe2d.SWEPhysics2D.from_case(c) for c in members]
    want = []
    for c, p in zip(members, phys):
        s = p.state(c.h, c.hu, c.hv)
        t = p.run(s, 0.0, n_steps, c.cfl)
        want.append((t, s.download(), p.masks(s)))
        s.free()
    ens = swe2d.Ensemble(phys)
    for m, c in enumerate(members):
        ens.upload(m, c.h, c.hu, c.hv)
    ens.run(1, members[0].cfl, graph=True, tune=True)
    times = ens.run(n_steps - 1, members[0].cfl, graph=True, tune=True)
    dry_members = 0
    for m, (t, x, masks) in enumerate(want):
        assert times[m] == t, (m, times[m], t)
        got = ens.download(m)
        for g, r in zip(got, x):
            assert np.array_equal(g, r), (m, np.abs(g - r).max())
        s = phys[m].state(*got)
        nd, es = phys[m].masks(s)
        s.free()
        assert np.array_equal(nd, masks[0]) and np.array_equal(es, masks[1]), m
        dry_members += int(nd.any())
    assert dry_members > 32  # wetting and drying is active in most members
    ens.close()


@pytest.fixture(scope="module")
def inversion_250k():
    inv = inversion.workload()   # 500 x 500 jittered quads over 50 km, N = 2, 20 stations every 10 of 2000 steps
    c = inv.case
    assert (c.n_elements, c.order, inv.n_steps) == (250_000, 2, 2000)
    phys = swe2d.SWEPhysics2D.from_case(c)
    inversion.observe(inv, phys)
    assert len(inv.dts) == 2000 and inv.obs.shape == (200, 20)
    x0 = phys.state(c.h, c.hu, c.hv)
    tape = phys.tape(20)
    J = tape.forward(x0, 0.0, inv.dts, inv.every, inv.obs)
    gbed, gman = tape.backward()
    stats = tape.stats()
    tape.close()
    x0.free()
    return inv, phys, J, gbed, gman, stats


def test_inversion_gradient_full_size_tape(inversion_250k):
    inv, phys, J, gbed, gman, (fsteps, vjps) = inversion_250k
    assert (fsteps, vjps) == (5980, 2000)   # 20 checkpoints over 2000 steps: the binomial schedule's count
    # the tape's misfit is the plain forward's; wet/dry is active at the end of the run
    c = inv.case
    J0, masks0 = _misfit_and_masks(phys, inv, inv.prior_bed, c.manning, (c.h, c.hu, c.hv))
    phys.set_bed(inv.prior_bed)
    assert abs(J - J0) <= 1e-12 * J, (J, J0)
    assert masks0[0].any() and masks0[1].any()   # wet/dry active at the end of the run
    assert np.abs(gbed).max() > 0.0 and np.abs(gman).max() > 0.0


@pytest.mark.parametrize("group", range(4))
def test_inversion_gradient_full_size_against_central_differences(inversion_250k, group):
    inv, phys, J, gbed, gman, _ = inversion_250k
    res = _differences(inv, phys, gbed, gman, range(8 * group, 8 * group + 8), range(32), 7000, "inversion-250k",
                       EPS_INVERSION)
    _check_set(res, f"inversion-250k directions {8 * group}-{8 * group + 7}")


@pytest.fixture(scope="module")
def inversion_ensemble():
    members = inversion.ensemble_workload(n_members=32)
    physics = []
    for inv in members:
        p = swe2d.SWEPhysics2D.from_case(inv.case)
        inversion.observe(inv, p)
        physics.append(p)
    grads = inversion.ensemble_gradients(members, physics, snapshots=20)
    return members, physics, grads


def test_inversion_ensemble_members(inversion_ensemble):
    members, physics, grads = inversion_ensemble
    ks = [m.case.n_elements for m in members]
    assert len(members) == 32 and min(ks) >= 19_000 and max(ks) <= 251_000
    assert {m.case.order for m in members} == {2, 3}
    for J, gbed, gman in grads:
        assert J > 0.0 and np.abs(gbed).max() > 0.0 and np.abs(gman).max() > 0.0


ENSEMBLE_ERRORS = {}


@pytest.mark.parametrize("group", range(4))
def test_inversion_ensemble_full_size_against_central_differences(inversion_ensemble, group):
    members, physics, grads = inversion_ensemble
    for m in range(8 * group, 8 * group + 8):
        J, gbed, gman = grads[m]
        res = _differences(members[m], physics[m], gbed, gman, range(8), range(8), 9000 + 101 * m, f"member {m}",
                           EPS_MEMBERS, allow_unmatched=True)
        ENSEMBLE_ERRORS[m] = [r[3] for r in res]
        assert np.nanmax(ENSEMBLE_ERRORS[m] + [0.0]) <= REL_GROSS, (m, ENSEMBLE_ERRORS[m])


def test_inversion_ensemble_error_distribution():
    errs = np.concatenate([ENSEMBLE_ERRORS[m] for m in sorted(ENSEMBLE_ERRORS)])
    assert errs.size == 256, errs.size
    unmatched = int(np.isnan(errs).sum())
    errs = errs[~np.isnan(errs)]
    med, p90, worst = np.median(errs), np.quantile(errs, 0.9), errs.max()
    print(f"inversion ensemble, 32 members x 8 directions: |fd - g| / s median {med:.2e}, 90 % {p90:.2e}, "
          f"worst {worst:.2e}; {unmatched} unmatched", flush=True)
    assert med <= REL and p90 <= REL_EACH and worst <= REL_GROSS and unmatched <= 4
