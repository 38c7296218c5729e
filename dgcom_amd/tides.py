"""Tidal astronomy for harmonic boundary tides: the equilibrium arguments V0,
the nodal corrections f and u of a constituent at an instant, their linear
drift through a run, and the constituent speeds (src/tides/astronomy.rs,
src/time/clock.rs).

Meeus' polynomial mean longitudes (Astronomical Algorithms, 2nd ed., 22, 25,
47) at a Julian date, Doodson's argument V0 = c_tau tau + c_s s + c_h h + c_p p
+ c_p1 p1 + offset, Schureman's (1958) closed-form f, u in the node longitude
N. The speeds are the time derivatives of the same combinations, so omega t
and V0 advance together. Only the constituents the runs here use are tabulated
(M2, S2); the arithmetic is the reference's operation for operation (Python
floats are IEEE doubles, `math.fmod` is Rust's `%`), so values agree with it to
the last bit except where a library cos/sin/atan2 differs by an ulp.
"""
from __future__ import annotations

import calendar
import math
from dataclasses import dataclass

DEG_TO_RAD = math.pi / 180.0
UNIX_EPOCH_JD = 2_440_587.5                     # clock.rs
DAY = 86_400.0

# Secular rates of Meeus' arguments, degrees per Julian century (astronomy.rs:62-72)
S_PER_CENTURY = 481_267.881_234_21              # Moon mean longitude L'
H_PER_CENTURY = 36_000.769_827_79               # Sun mean longitude L0
MP_PER_CENTURY = 477_198.867_505_5              # Moon mean anomaly M'
M_PER_CENTURY = 35_999.050_290_9                # Sun mean anomaly M
HOURS_PER_CENTURY = 36_525.0 * 24.0


def wrap_deg(x: float) -> float:
    """An angle in degrees wrapped to [0, 360) (astronomy.rs:79-82)."""
    r = math.fmod(x, 360.0)
    return r + 360.0 if r < 0.0 else r


@dataclass(frozen=True)
class AstronomicalArguments:
    tau: float  # mean lunar time 15 H_UT + h - s
    s: float    # Moon mean longitude
    h: float    # Sun mean longitude
    p: float    # lunar perigee
    n: float    # Moon's ascending node
    p1: float   # solar perigee

    @staticmethod
    def at_julian_date(jd_utc: float) -> "AstronomicalArguments":
        """Meeus (22.x, 25.2, 47.x) at a Julian date, UTC for TT (astronomy.rs:111-155)."""
        t = (jd_utc - 2_451_545.0) / 36_525.0
        t2 = t * t
        t3 = t2 * t
        t4 = t3 * t
        s = 218.316_447_7 + S_PER_CENTURY * t - 0.001_578_6 * t2 + t3 / 538_841.0 - t4 / 65_194_000.0
        h = 280.466_456_7 + H_PER_CENTURY * t + 0.000_303_2 * t2
        mp = 134.963_396_4 + MP_PER_CENTURY * t + 0.008_741_4 * t2 + t3 / 69_699.0 - t4 / 14_712_000.0
        m = 357.529_109_2 + M_PER_CENTURY * t - 0.000_153_6 * t2 + t3 / 24_490_000.0
        node = 125.044_547_9 - 1_934.136_289_1 * t + 0.002_075_4 * t2 + t3 / 467_441.0 - t4 / 60_616_000.0
        p = s - mp
        p1 = h - m
        day_fraction = (jd_utc + 0.5) % 1.0
        h_ut = day_fraction * 24.0
        tau = 15.0 * h_ut + h - s
        return AstronomicalArguments(wrap_deg(tau), wrap_deg(s), wrap_deg(h), wrap_deg(p), wrap_deg(node),
                                     wrap_deg(p1))


def julian_date(year, month, day, hour, minute, second) -> float:
    """Meeus (7.1), Gregorian calendar (astronomy.rs:172-189)."""
    y, m = (year - 1, month + 12) if month <= 2 else (year, month)
    a = math.floor(y / 100.0)
    b = 2.0 - a + math.floor(a / 4.0)
    day_fraction = (hour + minute / 60.0 + second / 3600.0) / 24.0
    return (math.floor(365.25 * (y + 4716.0)) + math.floor(30.6001 * (m + 1.0)) + day + b - 1524.5
            + day_fraction)


@dataclass(frozen=True)
class Doodson:
    c_tau: float
    c_s: float
    c_h: float
    c_p: float
    c_p1: float
    offset: float

    def v0(self, a: AstronomicalArguments) -> float:
        """The equilibrium argument in degrees (astronomy.rs:292-303)."""
        return wrap_deg(self.c_tau * a.tau + self.c_s * a.s + self.c_h * a.h + self.c_p * a.p + self.c_p1 * a.p1
                        + self.offset)

    def speed_deg_per_hour(self) -> float:
        """dV0/dt from the secular rates (astronomy.rs:306-314)."""
        s = S_PER_CENTURY / HOURS_PER_CENTURY
        h = H_PER_CENTURY / HOURS_PER_CENTURY
        p = (S_PER_CENTURY - MP_PER_CENTURY) / HOURS_PER_CENTURY
        p1 = (H_PER_CENTURY - M_PER_CENTURY) / HOURS_PER_CENTURY
        tau = 15.0 + h - s
        return self.c_tau * tau + self.c_s * s + self.c_h * h + self.c_p * p + self.c_p1 * p1


def nodal_m2(n_rad: float):
    """Schureman's f, u (degrees) of the M2 family (astronomy.rs:471-475)."""
    f = 1.0004 - 0.0373 * math.cos(n_rad) + 0.0002 * math.cos(2.0 * n_rad)
    u = -2.14 * math.sin(n_rad)
    return f, u


# name -> (Doodson argument, nodal family) (astronomy.rs:360-367)
CONSTITUENTS = {
    "M2": (Doodson(2.0, 0.0, 0.0, 0.0, 0.0, 0.0), "M2"),
    "S2": (Doodson(2.0, 2.0, -2.0, 0.0, 0.0, 0.0), None),
}


@dataclass(frozen=True)
class NodalCorrection:
    f: float
    u_deg: float
    v0_deg: float

    def phase_offset_rad(self) -> float:
        return (self.v0_deg + self.u_deg) * DEG_TO_RAD


def nodal_correction(name: str, a: AstronomicalArguments) -> NodalCorrection:
    """f, u and V0 of a constituent at astronomical arguments (astronomy.rs:577-584)."""
    doodson, kind = CONSTITUENTS[name.upper()]
    f, u = nodal_m2(a.n * DEG_TO_RAD) if kind == "M2" else (1.0, 0.0)
    return NodalCorrection(f, u, doodson.v0(a))


def angular_speed(name: str) -> float:
    """Degrees per mean solar hour (astronomy.rs:594-596)."""
    return CONSTITUENTS[name.upper()][0].speed_deg_per_hour()


def constituent_period(name: str) -> float:
    """Seconds, 360 / speed (astronomy.rs:601-603)."""
    return 360.0 / angular_speed(name) * 3600.0


@dataclass(frozen=True)
class NodalDrift:
    """f and u linear in the time since the epoch, V0 at the epoch (astronomy.rs:234-259)."""
    at_epoch: NodalCorrection
    f_rate: float
    u_rate_deg: float

    def at(self, t: float) -> NodalCorrection:
        return NodalCorrection(self.at_epoch.f + self.f_rate * t, self.at_epoch.u_deg + self.u_rate_deg * t,
                               self.at_epoch.v0_deg)


class ModelClock:
    """Simulation time t = seconds since an epoch (clock.rs)."""

    def __init__(self, epoch_unix: float):
        self.epoch_unix = float(epoch_unix)

    @staticmethod
    def at_datetime(year, month, day, hour=0, minute=0, second=0) -> "ModelClock":
        return ModelClock(calendar.timegm((year, month, day, hour, minute, second)))

    def julian_date(self, t: float) -> float:
        return UNIX_EPOCH_JD + (self.epoch_unix + t) / DAY

    def astronomical_arguments(self, t: float) -> AstronomicalArguments:
        return AstronomicalArguments.at_julian_date(self.julian_date(t))

    def nodal_drift(self, name: str) -> NodalDrift:
        """Central differences over +-1 day about the epoch (clock.rs:108-117)."""
        before = nodal_correction(name, self.astronomical_arguments(-DAY))
        at_epoch = nodal_correction(name, self.astronomical_arguments(0.0))
        after = nodal_correction(name, self.astronomical_arguments(DAY))
        return NodalDrift(at_epoch, (after.f - before.f) / (2.0 * DAY), (after.u_deg - before.u_deg) / (2.0 * DAY))
