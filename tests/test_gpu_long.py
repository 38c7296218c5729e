"""Whole-run acceptance of the HIP path against the CPU restatement of the
reference, through the C ABI:

- the idealised tidal estuary in full (100 x 20 quads, N = 2) and the large
  unstructured workload's physics on a 64 x 64 jittered mesh (N = 3, wet/dry
  active) over 1000 SSP-RK3 steps: every time step, elevation and transports
  bit for bit (so relative L2 = 0 <= 1e-11), identical integer wet/dry masks and
  clip counts;
- the full 1,000,000-element workload: its first steps bit for bit;
- mass conserved to 1e-13 relative over 1000 steps in a walled basin that wets
  and dries;
- lake at rest across shorelines held to 1e-14 m over 1000 steps.
"""
import numpy as np
import pytest

from dgcom_amd import cases, swe2d
import oracle as O

pytestmark = pytest.mark.gpu


def run_both(case, steps):
    orc = O.Oracle.from_case(case)
    phys = swe2d.SWEPhysics2D.from_case(case)
    h, hu, hv = case.h.copy(), case.hu.copy(), case.hv.copy()
    s = phys.state(h, hu, hv)
    te_g, dts_g = phys.run(s, 0.0, steps, case.cfl, record_dt=True)
    te_o, dts_o = orc.run(h, hu, hv, 0.0, steps, case.cfl, record_dt=True)
    return orc, phys, s, (h, hu, hv), (te_g, dts_g), (te_o, dts_o)


@pytest.mark.parametrize("name,make", [
    ("tidal_estuary_100x20_n2", lambda: cases.tidal_estuary()),
    ("unstructured_64x64_n3", lambda: cases.large_unstructured(n=64, length=6400.0, correlation=400.0, sd=25.0)),
])
def test_thousand_steps_bitwise(name, make):
    case = make()
    print(f"{name}: K={case.n_elements} N={case.order}", flush=True)
    orc, phys, s, ref, (te_g, dts_g), (te_o, dts_o) = run_both(case, 1000)
    assert np.array_equal(dts_g, dts_o)
    assert te_g == te_o
    got = s.download()
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    eta_g, eta_o = got[0] + case.bed, ref[0] + case.bed
    assert np.linalg.norm(eta_g - eta_o) <= 1e-11 * np.linalg.norm(eta_o)
    nd_g, es_g = phys.masks(s)
    nd_o, es_o = orc.masks(ref[0])
    assert np.array_equal(nd_g, nd_o) and np.array_equal(es_g, es_o)
    assert phys.negative_depth_clips() == orc.clips()
    print(f"{name}: t = {te_g:.3f} s, {int(es_g.sum())} subcell elements, {int(nd_g.sum())} dry nodes", flush=True)


def test_full_size_workload_first_steps_bitwise():
    case = cases.large_unstructured()
    print(f"full workload: K={case.n_elements}", flush=True)
    orc, phys, s, ref, (te_g, dts_g), (te_o, dts_o) = run_both(case, 3)
    assert np.array_equal(dts_g, dts_o)
    for g, r in zip(s.download(), ref):
        assert np.array_equal(g, r)
    nd_g, es_g = phys.masks(s)
    nd_o, es_o = orc.masks(ref[0])
    assert np.array_equal(nd_g, nd_o) and np.array_equal(es_g, es_o)


def test_mass_conserved_in_walled_basin_over_thousand_steps():
    case = cases.drying_beach(order=3, nx=32, ny=8, jitter=0.2, seed=5)
    case.mesh.tags[:] = cases.WALL
    rng = np.random.default_rng(5)
    h, hu, hv = cases.still_water(case.bed)
    # a sloshing surface over the beach: water runs up and down the shore
    x, _ = cases.node_coordinates(case.mesh, case.order)
    h = np.maximum(h + 0.3 * np.cos(np.pi * x.ravel() / 2000.0) * (h > 0), 0.0)
    phys = swe2d.SWEPhysics2D.from_case(case)
    orc = O.Oracle.from_case(case)
    mass_w = orc.geometry()["mass"]
    s = phys.state(h, hu, hv)
    m0 = np.sum(mass_w * h)
    phys.run(s, 0.0, 1000, case.cfl)
    hg = s.download()[0]
    nd, es = phys.masks(s)
    assert es.sum() > 0  # wetting and drying took place
    assert phys.negative_depth_clips() == 0
    assert abs(np.sum(mass_w * hg) - m0) / m0 <= 1e-13


def test_lake_at_rest_over_thousand_steps():
    case = cases.drying_beach(order=2, nx=32, ny=8, jitter=0.2, seed=9)
    bed = case.bed + np.random.default_rng(9).uniform(-0.5, 0.5, case.bed.size)  # rough, face-discontinuous
    h, hu, hv = cases.still_water(bed)
    phys = swe2d.SWEPhysics2D(case.order, case.mesh, bed, manning=case.manning, f0=case.f0)
    s = phys.state(h, hu, hv)
    phys.run(s, 0.0, 1000, case.cfl)
    hg, hug, hvg = s.download()
    wet = hg > 1e-3
    nd, es = phys.masks(s)
    assert es.sum() > 0
    assert np.abs((hg + bed)[wet]).max() <= 1e-14
    assert np.array_equal(hg <= 0.0, h <= 0.0)
    # momentum stays at round-off (the surface bar above is the 1e-14 m one)
    assert np.abs(hug).max() <= 1e-11 and np.abs(hvg).max() <= 1e-11
