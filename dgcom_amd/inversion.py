"""The bathymetry inversion workload of the adjoint: a truth bed and a prior
bed that differs from it by a smooth random field, stations, observations of
the truth run, and the station misfit of the prior's run.

Jittered quads over a square basin, N = 2, truth bed = -(mean depth + a seeded
Gaussian random field) with shorelines (wet/dry active), prior = truth + a
second seeded field of standard deviation `prior_sd`; a 0.5 m Gaussian hump of
water level at rest in the middle of the basin sets the waves going (at rest
the water level is flat whatever the bed, so observations need motion), an M2
tide on the west boundary; `n_stations` seeded stations in water, sampled after
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


def hump_state(case, bed, amplitude=0.5, radius_fraction=0.05):
    x, y = cases.node_coordinates(case.mesh, case.order)
    c = case.mesh.corners
    x0, x1 = c[:, :, 0].min(), c[:, :, 0].max()
    y0, y1 = c[:, :, 1].min(), c[:, :, 1].max()
    xc, yc = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    sig = radius_fraction * (x1 - x0)
    eta = amplitude * np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (2.0 * sig * sig)).ravel()
    h = np.maximum(eta - bed, 0.0)
    return h, np.zeros_like(h), np.zeros_like(h)


def workload(n=500, length=50_000.0, order=2, seed=4321, prior_sd=2.0, n_stations=20, every=10, n_steps=2000,
             mean_depth=40.0, sd=25.0, correlation=5000.0) -> Inversion:
    truth = cases.large_unstructured(n=n, length=length, order=order, seed=seed, n_steps=n_steps,
                                     correlation=correlation, mean_depth=mean_depth, sd=sd)
    x, y = cases.node_coordinates(truth.mesh, order)
    pert = cases.gaussian_random_field(x, y, (0.0, length, 0.0, length), prior_sd, correlation, seed + 7).ravel()
    prior_bed = truth.bed + pert
    st = stations.random_stations(truth, n_stations, seed + 11, min_depth=5.0)
    h, hu, hv = hump_state(truth, prior_bed)
    case = cases.Case("inversion_prior", order, truth.mesh, prior_bed, truth.manning, truth.f0, truth.tide, h, hu, hv,
                      cfl=0.5, scheme=0, n_steps=n_steps)
    return Inversion(case, truth.bed.copy(), prior_bed, st, every, n_steps)


def observe(inv: Inversion, phys):
    """Record the prior run's time steps, then the truth run's station levels at
    the same steps (from the same initial water level over the truth bed).
    `phys` is a SWEPhysics2D of inv.case; its bed is left at the prior."""
    c = inv.case
    phys.set_bed(inv.prior_bed)
    s = phys.state(c.h, c.hu, c.hv)
    _, inv.dts = phys.run(s, 0.0, inv.n_steps, c.cfl, record_dt=True)
    phys.set_bed(inv.truth_bed)
    phys.set_stations(inv.stations.elements, inv.stations.weights)
    s = phys.state(*hump_state(c, inv.truth_bed))
    obs, t = [], 0.0
    for i, dt in enumerate(inv.dts):
        phys.step(s, t, dt)
        t += dt
        if (i + 1) % inv.every == 0:
            obs.append(phys.sample_stations(s))
    inv.obs = np.array(obs)
    phys.set_bed(inv.prior_bed)
    return inv
