"""Local time stepping on the GPU (dg_begin_coarse_step / dg_stage_where /
dg_finish_where driven by dgcom_amd.multirate) against the restatement driven by
the same stepper: whole multirate runs bit for bit (levels, coarse steps and
states), on the reference's graded channel (levels 0-2, mass to rounding), its
island lake at rest (wet/dry, friction) and the tidal drying beach."""
import numpy as np
import pytest

from dgcom_amd import cases, multirate as M, swe2d
import oracle as O

pytestmark = pytest.mark.gpu


def _runs(c, t_end, cfl, friction, wet_dry, max_levels=8):
    nb = M.neighbours(c.mesh.links)
    orc = O.Oracle.from_case(c, friction=friction, wet_dry=wet_dry)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    t_o, n_o, st_o = M.run(orc, nb, x, 0.0, t_end, cfl, max_levels, c.scheme)
    phys = swe2d.SWEPhysics2D.from_case(c, friction=friction, wet_dry=wet_dry)
    local = phys.local_time_stepping()
    s = phys.state(c.h, c.hu, c.hv)
    t_g, n_g, st_g = M.run(local, nb, s, 0.0, t_end, cfl, max_levels, c.scheme)
    got = s.download()
    st_g.release(local)
    return (t_o, n_o, st_o, x, orc), (t_g, n_g, st_g, got, phys)


@pytest.mark.parametrize("scheme", [0, 1])
def test_graded_channel_equals_the_oracle(scheme):
    c = cases.graded_channel(scheme=scheme)
    (t_o, n_o, st_o, x, _), (t_g, n_g, st_g, got, _) = _runs(c, 400.0, 0.5, False, False)
    assert (t_g, n_g) == (t_o, n_o)
    assert st_g.stats == st_o.stats and st_g.stats.finest_level >= 2
    for a, b in zip(got, x):
        assert np.array_equal(a, b)
    m = cases.node_masses(c)
    assert abs(float(m @ got[0]) - float(m @ c.h)) / float(m @ c.h) < 1e-13


@pytest.mark.parametrize("scheme", [0, 1])
def test_island_lake_at_rest_equals_the_oracle(scheme):
    c = cases.graded_channel(island=True, friction=True, scheme=scheme, cfl=1.0)
    (t_o, n_o, st_o, x, orc), (t_g, n_g, st_g, got, phys) = _runs(c, 600.0, 1.0, True, True)
    assert (t_g, n_g) == (t_o, n_o) and st_g.stats.finest_level >= 2
    for a, b in zip(got, x):
        assert np.array_equal(a, b)
    assert phys.negative_depth_clips() == 0
    assert np.abs(got[0] - c.h).max() < 1e-10


def test_tidal_drying_beach_equals_the_oracle():
    c = cases.drying_beach(nx=24, ny=6)
    (t_o, n_o, st_o, x, orc), (t_g, n_g, st_g, got, phys) = _runs(c, 120.0, c.cfl, True, True)
    assert (t_g, n_g) == (t_o, n_o)
    assert st_g.stats == st_o.stats
    for a, b in zip(got, x):
        assert np.array_equal(a, b)
    nd_g, es_g = phys.masks(phys.state(*got))
    nd_o, es_o = orc.masks(x[0])
    assert np.array_equal(nd_g, nd_o) and np.array_equal(es_g, es_o) and es_g.any()
    assert phys.negative_depth_clips() == orc.clips()


@pytest.mark.parametrize("velocity", ["at_rest", "incoming_wave"])
def test_tidal_channel_lts_equals_the_oracle(velocity):
    """The reference's 3-level tidal channel gate: 25 h, 1,470 coarse steps
    (1,529 with the incoming wave), the state bit for bit."""
    c = cases.tidal_channel(velocity=velocity)
    nb = M.neighbours(c.mesh.links)
    orc = O.Oracle.from_case(c)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    _, n_o, _ = M.run(orc, nb, x, 0.0, 25 * 3600.0, c.cfl, 3, c.scheme, interval=1800.0)
    phys = swe2d.SWEPhysics2D.from_case(c)
    local = phys.local_time_stepping()
    s = phys.state(c.h, c.hu, c.hv)
    _, n_g, st = M.run(local, nb, s, 0.0, 25 * 3600.0, c.cfl, 3, c.scheme, interval=1800.0)
    assert n_g == n_o == {"at_rest": 1470, "incoming_wave": 1529}[velocity]
    for a, b in zip(s.download(), x):
        assert np.array_equal(a, b)
    assert phys.negative_depth_clips() == 0


# ---------------------------------------------------------------------------
# The library's own stepper (dg_multirate_*, C++) against the restatement
# driven by dgcom_amd.multirate
# ---------------------------------------------------------------------------
def _library_run(c, t_end, cfl, friction, wet_dry, max_levels=8, interval=None):
    phys = swe2d.SWEPhysics2D.from_case(c, friction=friction, wet_dry=wet_dry)
    mr = swe2d.Multirate(phys, max_levels, c.scheme)
    s = phys.state(c.h, c.hu, c.hv)
    t, n = mr.run(s, 0.0, t_end, cfl, interval=interval)
    got = s.download()
    stats = mr.stats()
    mr.close()
    return t, n, got, stats, phys


def _oracle_run(c, t_end, cfl, friction, wet_dry, max_levels=8, interval=None):
    orc = O.Oracle.from_case(c, friction=friction, wet_dry=wet_dry)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    t, n, st = M.run(orc, M.neighbours(c.mesh.links), x, 0.0, t_end, cfl, max_levels, c.scheme, interval=interval)
    return t, n, x, st, orc


@pytest.mark.parametrize("which", ["graded_rk3", "graded_rk43", "island_rk3", "island_rk43", "beach"])
def test_library_stepper_equals_the_oracle(which):
    if which.startswith("graded"):
        c = cases.graded_channel(scheme=0 if which.endswith("rk3") else 1)
        args = (400.0, 0.5, False, False)
    elif which.startswith("island"):
        c = cases.graded_channel(island=True, friction=True, scheme=0 if which.endswith("rk3") else 1, cfl=1.0)
        args = (600.0, 1.0, True, True)
    else:
        c = cases.drying_beach(nx=24, ny=6)
        args = (120.0, c.cfl, True, True)
    t_o, n_o, x, st, orc = _oracle_run(c, *args)
    t_g, n_g, got, stats, phys = _library_run(c, *args)
    assert (t_g, n_g) == (t_o, n_o)
    assert stats["element_evaluations"] == st.stats.element_evaluations
    assert stats["finest_level"] == st.stats.finest_level
    for a, b in zip(got, x):
        assert np.array_equal(a, b)
    assert phys.negative_depth_clips() == orc.clips()


@pytest.mark.parametrize("velocity", ["at_rest", "incoming_wave"])
def test_library_stepper_tidal_channel(velocity):
    c = cases.tidal_channel(velocity=velocity)
    t_o, n_o, x, st, orc = _oracle_run(c, 25 * 3600.0, c.cfl, True, True, 3, 1800.0)
    t_g, n_g, got, stats, phys = _library_run(c, 25 * 3600.0, c.cfl, True, True, 3, 1800.0)
    assert n_g == n_o == {"at_rest": 1470, "incoming_wave": 1529}[velocity]
    for a, b in zip(got, x):
        assert np.array_equal(a, b)
