"""Parity of the HIP path (through the C ABI) with the CPU restatement of the
reference on the same seeded inputs.

Bars: the product build (no multiply-add contraction) is compared bit for bit:
every RHS, stage, time step and run, wet/dry masks and clip counts included.
The FMA variant (multiply-adds contracted) is held to 1e-12 of the field
maximum per RHS and stage; over a run its rounding differences grow (to ~1e-11
relative L2 of the elevation after 300 steps of the wet estuary, and faster
where near-dry velocities set the time step), so its run is held to 1e-9 on
the wet estuary only.
"""
import math

import numpy as np
import pytest

from dgcom_amd import cases
from dgcom_amd import swe2d
import oracle as O
from cases_for_tests import CASES, perturbed, small_unstructured

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def rel_l2(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("fma", [False, True], ids=["exact", "fma"])
def test_rhs_parity(name, make, fma):
    case = make()
    print(f"case {name} K={case.n_elements} N={case.order} fma={fma}", flush=True)
    orc = O.Oracle.from_case(case)
    phys = swe2d.SWEPhysics2D.from_case(case, fma=fma)
    for t in (0.0, 1800.0):
        h, hu, hv = perturbed(case, seed=int(t))
        ref = orc.rhs(h, hu, hv, t)
        s = phys.state(h, hu, hv)
        out = phys.state()
        phys.compute_rhs_into(s, t, out)
        got = out.download()
        for g, r in zip(got, ref):
            if not fma:
                bad = np.flatnonzero(g != r)
                assert bad.size == 0, f"{bad.size} values differ, first {bad[:5]}: {g[bad[:5]]} vs {r[bad[:5]]}"
            else:
                assert rel_max(g, r) <= 1e-12


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_dt_parity(name, make):
    case = make()
    orc = O.Oracle.from_case(case)
    phys = swe2d.SWEPhysics2D.from_case(case, fma=False)
    h, hu, hv = perturbed(case, seed=3)
    s = phys.state(h, hu, hv)
    got = phys.element_dt(s, case.cfl)
    ref = orc.element_dt(h, hu, hv, case.cfl)
    assert np.array_equal(got, ref)
    assert phys.compute_dt_ssp(s, case.cfl) == orc.dt(h, hu, hv, case.cfl)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("fma", [False, True], ids=["exact", "fma"])
def test_stage_parity(name, make, fma):
    case = make()
    orc = O.Oracle.from_case(case)
    phys = swe2d.SWEPhysics2D.from_case(case, fma=fma)
    h, hu, hv = perturbed(case, seed=5)
    b = [a.copy() for a in perturbed(case, seed=6)]
    dt = orc.dt(h, hu, hv, case.cfl)
    for (a_, b_, c_, last) in ((1.0, 0.0, dt, False), (0.75, 0.25, 0.25 * dt, False), (1 / 3, 2 / 3, 2 / 3 * dt, True)):
        ref = orc.stage(b, (h, hu, hv), a_, b_, c_, 100.0, last)
        sb, si, so = phys.state(*b), phys.state(h, hu, hv), phys.state()
        phys.stage_where(sb, si, so, a_, b_, c_, 100.0)
        got = so.download()
        for g, r in zip(got, ref):
            if not fma:
                assert np.array_equal(g, r)
            else:
                assert rel_max(g, r) <= 1e-12


@pytest.mark.parametrize("name,make,steps", [
    ("estuary_n2", lambda: cases.tidal_estuary(nx=20, ny=4), 300),
    ("beach_n2_jitter", lambda: cases.drying_beach(order=2, jitter=0.2), 200),
    ("beach_n3", lambda: cases.drying_beach(order=3), 200),
    ("unstructured_n3", lambda: small_unstructured(3), 200),
    ("unstructured_n2", lambda: small_unstructured(2), 200),
], ids=["estuary_n2", "beach_n2_jitter", "beach_n3", "unstructured_n3", "unstructured_n2"])
@pytest.mark.parametrize("fma", [False, True], ids=["exact", "fma"])
def test_run_parity(name, make, steps, fma):
    if fma and name != "estuary_n2":
        pytest.skip("FMA runs with wet/dry time steps drift apart; parity is held on the wet estuary")
    case = make()
    print(f"run {name} K={case.n_elements} N={case.order} steps={steps} fma={fma}", flush=True)
    orc = O.Oracle.from_case(case)
    phys = swe2d.SWEPhysics2D.from_case(case, fma=fma)
    h, hu, hv = case.h.copy(), case.hu.copy(), case.hv.copy()
    s = phys.state(h, hu, hv)
    te_g, dts_g = phys.run(s, 0.0, steps, case.cfl, record_dt=True)
    te_o, dts_o = orc.run(h, hu, hv, 0.0, steps, case.cfl, record_dt=True)
    got = s.download()
    nd_g, es_g = phys.masks(s)
    nd_o, es_o = orc.masks(h)
    assert np.array_equal(nd_g, nd_o) and np.array_equal(es_g, es_o)
    assert phys.negative_depth_clips() == orc.clips()
    if not fma:
        assert np.array_equal(dts_g, dts_o)
        for g, r in zip(got, (h, hu, hv)):
            assert np.array_equal(g, r)
    else:
        assert abs(te_g - te_o) <= 1e-12 * te_o
        eta_g = got[0] + case.bed
        eta_o = h + case.bed
        assert rel_l2(eta_g, eta_o) <= 1e-9
        assert rel_l2(got[1], hu) <= 1e-9 and rel_l2(got[2], hv) <= 1e-9
