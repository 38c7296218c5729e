"""The sources given per node on the GPU against the restatement, global steps: Manning n per node
(SpatiallyVaryingManning2D, friction.rs:463-553) and the beta-plane (coriolis.rs:101-121,157-169), bit for
bit over whole runs with wet/dry switching; a per-node field that is uniform over each element runs like the
per-element n; the adjoint and the ensembles refuse both by name; and one context used from two host threads at
once is serialised (each thread's run equals its run alone)."""
import threading

import numpy as np
import pytest

from dgcom_amd import cases, swe2d
from cases_for_tests import perturbed
import oracle as O

pytestmark = pytest.mark.gpu


def _run_both(c, setup, n_steps=40):
    orc = O.Oracle.from_case(c)
    phys = swe2d.SWEPhysics2D.from_case(c)
    setup(orc, phys)
    h, hu, hv = c.h.copy(), c.hu.copy(), c.hv.copy()
    t_cpu = orc.run(h, hu, hv, 0.0, n_steps, c.cfl)
    s = phys.state(c.h, c.hu, c.hv)
    t_gpu = phys.run(s, 0.0, n_steps, c.cfl)
    return (t_cpu, (h, hu, hv)), (t_gpu, s.download()), orc, phys


@pytest.mark.parametrize("order", [1, 2, 3])
def test_manning_per_node_equals_the_restatement(order):
    c = cases.drying_beach(order=order, nx=16, ny=5, jitter=0.2)
    n_node = np.repeat(c.manning, c.nn) * np.random.default_rng(order).uniform(0.3, 2.0, c.h.size)

    def setup(orc, phys):
        orc.set_manning_nodal(n_node)
        phys.set_manning_nodal(n_node)

    (tc, xc), (tg, xg), orc, phys = _run_both(c, setup)
    assert tc == tg
    for a, b in zip(xg, xc):
        assert np.array_equal(a, b)
    assert orc.masks(xc[0])[1].sum() > 0   # subcell elements: the wet/dry path ran


def test_manning_per_node_uniform_in_each_element_is_the_per_element_n():
    c = cases.drying_beach(order=2, nx=16, ny=5, jitter=0.2)
    phys = swe2d.SWEPhysics2D.from_case(c)
    s0 = phys.state(c.h, c.hu, c.hv)
    t0 = phys.run(s0, 0.0, 40, c.cfl)
    phys.set_manning_nodal(np.repeat(c.manning, c.nn))
    s1 = phys.state(c.h, c.hu, c.hv)
    t1 = phys.run(s1, 0.0, 40, c.cfl)
    assert t0 == t1
    for a, b in zip(s0.download(), s1.download()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_beta_plane_equals_the_restatement(order):
    c = cases.drying_beach(order=order, nx=16, ny=5, jitter=0.2)

    def setup(orc, phys):
        orc.set_beta(1e-8, 250.0)
        phys.set_coriolis_beta(1e-8, 250.0)

    (tc, xc), (tg, xg), orc, phys = _run_both(c, setup)
    assert tc == tg
    for a, b in zip(xg, xc):
        assert np.array_equal(a, b)
    # and the plane does act: the run differs from the f-plane's
    phys.set_coriolis_beta(0.0, 0.0)
    s = phys.state(c.h, c.hu, c.hv)
    phys.run(s, 0.0, 40, c.cfl)
    assert not np.array_equal(s.download()[1], xg[1])


def test_adjoint_and_ensembles_refuse_the_per_node_sources():
    c = cases.drying_beach(order=2, nx=8, ny=3)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_manning_nodal(np.repeat(c.manning, c.nn))
    x = phys.state(c.h, c.hu, c.hv)
    w = phys.state(c.h, c.hu, c.hv)
    g = phys.state()
    with pytest.raises(RuntimeError, match="Manning n per node"):
        phys.stage_vjp(x, x, 1.0, 0.0, 1.0, 0.0, w, None, g)
    with pytest.raises(RuntimeError, match="per node or a beta-plane"):
        swe2d.Ensemble([phys])
    phys.set_manning_nodal(None)
    phys.set_coriolis_beta(1e-9, 0.0)
    with pytest.raises(RuntimeError, match="beta-plane"):
        phys.stage_vjp(x, x, 1.0, 0.0, 1.0, 0.0, w, None, g)


def test_one_context_from_two_threads_is_serialised():
    c = cases.drying_beach(order=2, nx=24, ny=6, jitter=0.1)
    starts = [perturbed(c, seed=i + 1, amp=0.02) for i in range(2)]
    phys = swe2d.SWEPhysics2D.from_case(c)
    alone = []
    for st in starts:
        s = phys.state(*st)
        t = 0.0
        for _ in range(30):
            dt = phys.compute_dt_ssp(s, c.cfl)
            phys.step(s, t, dt)
            t += dt
        alone.append((t, s.download()))
    states = [phys.state(*st) for st in starts]
    results = [None, None]
    errors = []

    def worker(i):
        try:
            s, t = states[i], 0.0
            for _ in range(30):
                dt = phys.compute_dt_ssp(s, c.cfl)
                phys.step(s, t, dt)
                t += dt
            results[i] = (t, s.download())
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for (ta, xa), (tb, xb) in zip(alone, results):
        assert ta == tb
        for a, b in zip(xa, xb):
            assert np.array_equal(a, b)
