"""The tidal astronomy restatement against the reference's own known values
(src/tides/astronomy.rs:615-760, src/time/clock.rs:143-220): Meeus' textbook
longitudes at J2000.0 and his Julian-date examples, the constituents' periods
and tabulated speeds, S2's solar argument, M2's nodal extremes, and the nodal
drift's accuracy over a month and a year."""
import math

import pytest

from dgcom_amd import tides as T


def test_mean_longitudes_at_j2000():
    a = T.AstronomicalArguments.at_julian_date(2_451_545.0)
    assert abs(a.s - 218.316_447_7) < 1e-4
    assert abs(a.h - 280.466_456_7) < 1e-4
    assert abs(a.p - 83.353_051_3) < 1e-3
    assert abs(a.n - 125.044_547_9) < 1e-4
    assert abs(a.p1 - 282.937_347_5) < 1e-3
    assert abs(a.tau - T.wrap_deg(15.0 * 12.0 + 280.466_456_7 - 218.316_447_7)) < 1e-3


def test_julian_date_reference_values():
    assert abs(T.julian_date(1957, 10, 4, 19, 26, 24.0) - 2_436_116.31) < 1e-2   # Meeus example 7.a
    assert abs(T.julian_date(2000, 1, 1, 12, 0, 0.0) - 2_451_545.0) < 1e-6


@pytest.mark.parametrize("name,period_h,speed", [("M2", 12.420_601_2, 28.984_104_2), ("S2", 12.0, 30.0)])
def test_periods_and_speeds(name, period_h, speed):
    assert abs(T.angular_speed(name) - speed) < 1e-6
    assert abs(T.constituent_period(name) / 3600.0 - period_h) < 1e-7
    # the rate of V0 over one day is the speed
    a0 = T.AstronomicalArguments.at_julian_date(2_451_545.0)
    a1 = T.AstronomicalArguments.at_julian_date(2_451_546.0)
    raw = T.nodal_correction(name, a1).v0_deg - T.nodal_correction(name, a0).v0_deg
    rate = 360.0 / period_h * 24.0
    turns = round((rate - raw) / 360.0)
    recovered = 360.0 / (raw + 360.0 * turns) * 24.0
    assert abs(recovered - period_h) < 1e-4


def test_s2_period_is_twelve_hours():
    assert abs(T.constituent_period("S2") - 12.0 * 3600.0) < 1e-9


def test_s2_argument_is_solar_time():
    a = T.AstronomicalArguments.at_julian_date(T.julian_date(2020, 6, 15, 0, 0, 0.0))
    v0 = T.nodal_correction("S2", a).v0_deg
    assert v0 < 1e-6 or abs(v0 - 360.0) < 1e-6
    a6 = T.AstronomicalArguments.at_julian_date(T.julian_date(2020, 6, 15, 6, 0, 0.0))
    assert abs(T.nodal_correction("S2", a6).v0_deg - 180.0) < 1e-6
    c = T.nodal_correction("S2", a)
    assert c.f == 1.0 and c.u_deg == 0.0


def test_m2_nodal_extremes():
    f0, u0 = T.nodal_m2(0.0)
    f180, u180 = T.nodal_m2(math.pi)
    assert abs(f0 - 0.9633) < 1e-3 and abs(u0) < 1e-12
    assert abs(f180 - 1.0379) < 1e-3 and abs(u180) < 1e-12


def test_nodal_drift_follows_f_and_u_through_a_run():
    clock = T.ModelClock.at_datetime(2025, 6, 15)
    month, year = 30.0 * 86_400.0, 365.25 * 86_400.0
    drift = clock.nodal_drift("M2")
    exact0 = T.nodal_correction("M2", clock.astronomical_arguments(0.0))
    assert drift.at(0.0) == exact0
    for t, du in ((month, 1e-3), (year, 0.13)):
        exact = T.nodal_correction("M2", clock.astronomical_arguments(t))
        lin = drift.at(t)
        assert lin.v0_deg == exact0.v0_deg
        assert abs(lin.u_deg - exact.u_deg) < du
        assert abs(lin.f - exact.f) < 2.5e-3
    assert clock.nodal_drift("S2").u_rate_deg == 0.0


def test_model_clock_epoch():
    clock = T.ModelClock.at_datetime(2025, 6, 1)
    assert clock.epoch_unix == 1_748_736_000.0
