"""The bathymetry inversion workload of the adjoint: a truth bed and a prior
bed that differs from it by a smooth random field, stations, observations of
the truth run, and the station misfit of the prior's run; and an ensemble of
such inversions over uneven meshes, whose members' gradients are taken in one
sequence of launches.

Jittered quads over a square basin, N = 2, truth bed = -(mean depth + a seeded
Gaussian random field) with shorelines (wet/dry active), prior = truth + a
second seeded field of standard deviation `prior_sd`; a 0.5 m Gaussian hump of
water level in the middle of the basin sets the waves going (at rest the water
level is flat whatever the bed, so observations need motion), with a seeded
1 mm / 1 cm/s jitter of level and velocity per node, an M2 tide on the west
boundary; `n_stations` seeded stations in water, sampled after
every `every`-th of `n_steps` SSP-RK3 steps. The time steps are those of the
prior's adaptive run, frozen for the misfit and its gradient.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import cases, stations


@dataclass
class Inversion:
    case: cases.Case          # the physics with the prior bed, and the initial state
    truth_bed: np.ndarray
    prior_bed: np.ndarray
    stations: stations.Stations
    every: int
    n_steps: int
    dts: np.ndarray = None
    obs: np.ndarray = None
    eta0: np.ndarray = None   # the initial water level and velocity, shared by the truth and prior runs
    u0: np.ndarray = None
    v0: np.ndarray = None


def hump_level(case, amplitude=0.5, radius_fraction=0.05):
    """A Gaussian hump of water level in the middle of the domain (0 elsewhere)."""
    x, y = cases.node_coordinates(case.mesh, case.order)
    c = case.mesh.corners
    x0, x1 = c[:, :, 0].min(), c[:, :, 0].max()
    y0, y1 = c[:, :, 1].min(), c[:, :, 1].max()
    xc, yc = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    sig = radius_fraction * (x1 - x0)
    return amplitude * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (2.0 * sig * sig)).ravel()


def initial_state(bed, eta0, u0, v0, wet_depth=0.01):
    """Depth max(eta0 - B, 0); momentum h (u0, v0) where the depth exceeds wet_depth."""
    h = np.maximum(eta0 - bed, 0.0)
    wet = h > wet_depth
    return h, np.where(wet, h * u0, 0.0), np.where(wet, h * v0, 0.0)


def workload(n=500, length=50_000.0, order=2, seed=4321, prior_sd=2.0, n_stations=20, every=10, n_steps=2000,
             mean_depth=40.0, sd=25.0, correlation=5000.0) -> Inversion:
    truth = cases.large_unstructured(n=n, length=length, order=order, seed=seed, n_steps=n_steps,
                                     correlation=correlation, mean_depth=mean_depth, sd=sd)
    x, y = cases.node_coordinates(truth.mesh, order)
    pert = cases.gaussian_random_field(x, y, (0.0, length, 0.0, length), prior_sd, correlation, seed + 7).ravel()
    prior_bed = truth.bed + pert
    st = stations.random_stations(truth, n_stations, seed + 11, min_depth=5.0)
    # The initial water level is the hump plus a 1 mm seeded jitter per node, the
    # initial velocity a seeded ~1 cm/s per node: at rest with the depth
    # continuous across faces, the HLL wave speeds tie at every face
    # (min(u_l - c_l, u_roe - c_roe) with equal arguments), the first stage is
    # not differentiable there and the gradient would be one of its one-sided
    # derivatives.
    noise = np.random.default_rng(seed + 13)
    eta0 = hump_level(truth) + 1e-3 * noise.standard_normal(truth.h.size)
    u0 = 0.01 * noise.standard_normal(truth.h.size)
    v0 = 0.01 * noise.standard_normal(truth.h.size)
    h, hu, hv = initial_state(prior_bed, eta0, u0, v0)
    case = cases.Case("inversion_prior", order, truth.mesh, prior_bed, truth.manning, truth.f0, truth.tide, h, hu, hv,
                      cfl=0.5, scheme=0, n_steps=n_steps)
    return Inversion(case, truth.bed.copy(), prior_bed, st, every, n_steps, eta0=eta0, u0=u0, v0=v0)


def observe(inv: Inversion, phys):
    """Record the prior run's time steps, then the truth run's station levels at
    the same steps (from the same initial water level and velocity over the truth
    bed). `phys` is a SWEPhysics2D of inv.case; its bed is left at the prior."""
    c = inv.case
    phys.set_bed(inv.prior_bed)
    s = phys.state(c.h, c.hu, c.hv)
    _, inv.dts = phys.run(s, 0.0, inv.n_steps, c.cfl, record_dt=True)
    phys.set_bed(inv.truth_bed)
    phys.set_stations(inv.stations.elements, inv.stations.weights)
    s = phys.state(*initial_state(inv.truth_bed, inv.eta0, inv.u0, inv.v0))
    obs, t = [], 0.0
    for i, dt in enumerate(inv.dts):
        phys.step(s, t, dt)
        t += dt
        if (i + 1) % inv.every == 0:
            obs.append(phys.sample_stations(s))
    inv.obs = np.array(obs)
    phys.set_bed(inv.prior_bed)
    return inv


def ensemble_workload(n_members=32, k_lo=20_000, k_hi=250_000, orders=(2, 3), n_stations=12, every=10, n_steps=1000,
                      seed=2024, prior_sd=2.0, cell=100.0, correlation=5000.0, mean_depth=30.0) -> list:
    """Members of an inversion ensemble, seeded: K log-uniform in [k_lo, k_hi]
    jittered quads of `cell` metres, N uniform over `orders`, each member with its
    own GRF truth bed (mean `mean_depth`, sd 25 m: wet/dry active), its own prior (truth +
    GRF of sd `prior_sd`), `n_stations` stations in water and its own hump of water
    level; `every`-th of `n_steps` SSP-RK3 steps sampled. Observations are made by
    `observe` (they need a GPU)."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_members):
        K = float(np.exp(rng.uniform(np.log(k_lo), np.log(k_hi))))
        n = max(4, int(round(np.sqrt(K))))
        order = int(orders[int(rng.integers(0, len(orders)))])
        # the member's bed is drawn again (next seed) until at least 1 % of it is
        # land, so that wetting and drying are active in every member
        for attempt in range(50):
            inv = workload(n=n, length=cell * n, order=order, seed=seed + 1000 + 31 * m + 100_003 * attempt,
                           prior_sd=prior_sd, n_stations=n_stations, every=every, n_steps=n_steps,
                           correlation=correlation, mean_depth=mean_depth)
            if float((inv.prior_bed > 0.0).mean()) >= 0.01:
                break
        else:
            raise RuntimeError(f"member {m}: no bed with land in 50 draws (mean_depth too large?)")
        out.append(inv)
    return out


def ensemble_gradients(members, physics, snapshots, keep_stages=False):
    """Every member's station misfit and its gradients with respect to the nodal
    bed and the per-element Manning n, from its own binomial tape of `snapshots`
    states (with `keep_stages`, also its step records: no reverse step recomputes
    its stages): all the members' forward sweeps, then all their reverse sweeps, on
    the library's stream. `members` are observed Inversions, `physics` their
    SWEPhysics2D (bed at the prior, stations set). Returns [(J, dJ/dB, dJ/dn)]."""
    tapes, Js = [], []
    for inv, phys in zip(members, physics):
        c = inv.case
        tape = phys.tape(snapshots)
        if keep_stages:
            tape.keep_stages()
        Js.append(tape.forward(phys.state(c.h, c.hu, c.hv), 0.0, inv.dts, inv.every, inv.obs))
        tapes.append(tape)
    out = []
    for J, tape in zip(Js, tapes):
        gbed, gman = tape.backward()
        out.append((J, gbed, gman))
        tape.close()
    return out
