"""Ensembles through the C ABI: members of orders 1-3 and different meshes,
forcings and time steps advanced together equal, bit for bit, each member's
own run (times included) and the restatement's adaptive run of that member
(state, time and wet/dry masks), with the stages launched directly, replayed
from HIP graphs, and after the launch-shape tuner."""
import numpy as np
import pytest

import oracle as O
from dgcom_amd import cases, swe2d
from cases_for_tests import perturbed, small_unstructured

pytestmark = pytest.mark.gpu


def _members():
    cs = [cases.drying_beach(order=1, nx=10, ny=4), cases.tidal_estuary(nx=20, ny=4),
          small_unstructured(3), cases.drying_beach(order=3, nx=8, ny=3, jitter=0.2), small_unstructured(2),
          cases.drying_beach(order=2, nx=12, ny=5, jitter=0.1)]
    states = [perturbed(c, seed=i + 1, amp=0.02) for i, c in enumerate(cs)]
    return cs, states


@pytest.mark.parametrize("graph,tune", [(False, False), (True, False), (True, True)])
def test_ensemble_equals_each_members_own_run(graph, tune):
    cs, states = _members()
    n_steps = 12
    want, clips = [], []
    for c, st in zip(cs, states):
        phys = swe2d.SWEPhysics2D.from_case(c)
        s = phys.state(*st)
        t = phys.run(s, 0.0, n_steps, c.cfl)
        want.append((t, s.download()))
        clips.append(phys.negative_depth_clips())
    phys = [swe2d.SWEPhysics2D.from_case(c) for c in cs]
    ens = swe2d.Ensemble(phys)
    for m, st in enumerate(states):
        ens.upload(m, *st)
    ens.run(5, cs[0].cfl, graph=graph, tune=tune)
    times = ens.run(n_steps - 5, cs[0].cfl, graph=graph, tune=tune)
    for m, (t, x) in enumerate(want):
        got = ens.download(m)
        assert times[m] == t, (m, times[m], t)
        assert phys[m].negative_depth_clips() == clips[m]   # the tuner's timing launches count none
        for g, r in zip(got, x):
            assert np.array_equal(g, r), (m, np.abs(g - r).max())
    # the restatement's own run of each member
    for m, (c, st) in enumerate(zip(cs, states)):
        orc = O.Oracle.from_case(c)
        h, hu, hv = (np.array(a, dtype=np.float64) for a in st)
        t_o = orc.run(h, hu, hv, 0.0, n_steps, c.cfl)
        assert times[m] == t_o, (m, times[m], t_o)
        got = ens.download(m)
        for g, r in zip(got, (h, hu, hv)):
            assert np.array_equal(g, r), (m, np.abs(g - r).max())
        nd_g, es_g = phys[m].masks(phys[m].state(*got))
        nd_o, es_o = orc.masks(h)
        assert np.array_equal(nd_g, nd_o) and np.array_equal(es_g, es_o)
    if tune:
        for o in (1, 2, 3):
            cfg = ens.tuning(o)
            assert cfg["subcell_groups"] > 0
