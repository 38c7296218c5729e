"""The reference's atlas case (`cases/tidal_channel/atlas.toml`, committed
fixtures with its `boundary_tides.txt`) on the GPU against the restatement:
the 25 h run through the case runner takes the reference's 5,787 steps with
every station record and the final state equal to the restatement's bit for
bit; with local time stepping at 3 levels, 1,504 coarse steps (notes/gpu.md:192),
through the host stepper and through the library's own, the state bit for bit."""
import os

import numpy as np
import pytest

import oracle as O
from dgcom_amd import case_file as CF, multirate as M, swe2d

pytestmark = pytest.mark.gpu

CASE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tidal_channel_atlas.toml")


class _OracleRunner:
    def __init__(self, case):
        self.orc = O.Oracle.from_case(case)

    def set_boundary_tides(self, bt):
        self.orc.set_boundary_tides(bt)

    def state(self, h, hu, hv):
        from test_case_file import _HostState
        return _HostState([np.array(a, dtype=np.float64) for a in (h, hu, hv)])

    def compute_dt_ssp(self, s, cfl, scheme):
        return self.orc.dt(*s.x, cfl, scheme)

    def step(self, s, t, dt, scheme):
        self.orc.run_dts(*s.x, t, [dt], scheme)

    def negative_depth_clips(self):
        return self.orc.clips()


def test_the_atlas_case_equals_the_restatement():
    run = CF.load(CASE)
    gpu = CF.run_case(run)
    ref = CF.run_case(CF.load(CASE), physics=_OracleRunner(run.case))
    assert gpu.steps == ref.steps == 5787 and gpu.clips == ref.clips == 0
    assert gpu.records == ref.records and len(gpu.records) == 153
    for a, b in zip(gpu.state, ref.state):
        assert np.array_equal(a, b)


def test_the_atlas_case_with_local_time_stepping():
    run = CF.load(CASE)
    c = run.case
    nb = M.neighbours(c.mesh.links)
    orc = O.Oracle.from_case(c)
    orc.set_boundary_tides(run.boundary_tides)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    _, n_o, _ = M.run(orc, nb, x, 0.0, 25 * 3600.0, c.cfl, 3, c.scheme, interval=1800.0)
    phys = swe2d.SWEPhysics2D.from_case(c)
    phys.set_boundary_tides(run.boundary_tides)
    local = phys.local_time_stepping()
    s = phys.state(c.h, c.hu, c.hv)
    _, n_g, st = M.run(local, nb, s, 0.0, 25 * 3600.0, c.cfl, 3, c.scheme, interval=1800.0)
    assert n_g == n_o == 1504
    for a, b in zip(s.download(), x):
        assert np.array_equal(a, b)
    st.release(local)
    assert phys.negative_depth_clips() == 0
    # the case runner with the library's own stepper
    rep = CF.run_case(CF.load(CASE, ["numerics.local_time_stepping_levels=3"]))
    assert rep.steps == 1504 and rep.clips == 0 and len(rep.records) == 153
