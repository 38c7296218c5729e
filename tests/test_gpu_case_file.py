"""The case runner on the GPU: the reference's tidal channel case file run as
its runner runs it (5,765 steps, 51 records per station, equal to the runner
over the restatement bit for bit), stopped at 12.5 h into a DGRST2D restart and
resumed to the same end bit for bit, and with local time stepping at 3 levels
(1,470 coarse steps)."""
import os

import numpy as np
import pytest

from dgcom_amd import case_file as CF
from test_case_file import CASE, OraclePhysics

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def straight():
    return CF.run_case(CF.load(CASE))


def test_the_case_equals_the_restatement(straight, tmp_path):
    run = CF.load(CASE)
    ref = CF.run_case(run, physics=OraclePhysics(run.case))
    assert straight.steps == ref.steps == 5765 and straight.clips == 0
    assert straight.records == ref.records
    for a, b in zip(straight.state, ref.state):
        assert np.array_equal(a, b)
    CF.write_stations(str(tmp_path / "stations.csv"), straight.records)
    lines = open(tmp_path / "stations.csv").read().splitlines()
    assert lines[0] == "station,time_s,eta_m,u_m_s,v_m_s,depth_m" and len(lines) == 1 + 153
    assert lines[1].startswith("mouth,0,")


def test_a_restart_resumes_bit_for_bit(straight, tmp_path):
    path = str(tmp_path / "half.dgrst")
    half = CF.run_case(CF.load(CASE), stop_at=12.5 * 3600.0, restart_to=path)
    assert half.t == 45_000.0
    rest = CF.run_case(CF.load(CASE), restart_from=path)
    assert rest.t == 90_000.0 and rest.steps == straight.steps
    assert rest.records == straight.records
    for a, b in zip(rest.state, straight.state):
        assert np.array_equal(a, b)


def test_the_case_with_local_time_stepping():
    rep = CF.run_case(CF.load(CASE, ["numerics.local_time_stepping_levels=3"]))
    assert rep.steps == 1470 and rep.clips == 0
    assert len(rep.records) == 153
