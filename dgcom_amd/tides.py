"""Tidal astronomy for harmonic boundary tides: the equilibrium arguments V0,
the nodal corrections f and u of a constituent at an instant, their linear
drift through a run, and the constituent speeds (src/tides/astronomy.rs,
src/time/clock.rs).

Meeus' polynomial mean longitudes (Astronomical Algorithms, 2nd ed., 22, 25,
47) at a Julian date, Doodson's argument V0 = c_tau tau + c_s s + c_h h + c_p p
+ c_p1 p1 + offset, Schureman's (1958) closed-form f, u in the node longitude
N (and, for L2, the lunar perigee p). The speeds are the time derivatives of
the same combinations, so omega t and V0 advance together. The constituent set
is the reference's 21 (astronomy.rs:343-458), with the minor constituents'
inference by linear admittance between two references (src/tides/admittance.rs).
The arithmetic is the reference's operation for operation (Python floats are
IEEE doubles, `math.fmod` is Rust's `%`, the trigonometry and `hypot` the
platform libm's that Rust calls too), so values agree with it to the last bit.
"""
from __future__ import annotations

import calendar
import ctypes
import ctypes.util
import math
from dataclasses import dataclass

# The C library's hypot (Rust's f64::hypot calls it; Python's math.hypot is
# CPython's own algorithm, which may round differently)
_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.hypot.restype = ctypes.c_double
_libm.hypot.argtypes = [ctypes.c_double, ctypes.c_double]


_libm.sincos.restype = None
_libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]


def hypot(x: float, y: float) -> float:
    return _libm.hypot(x, y)


def sincos(x: float):
    """(sin x, cos x) by the C library's sincos. The reference's `sin_cos()`, and a sine and
    cosine of one argument in one block, compile to this one call (LLVM merges the pair), and
    glibc's sincos differs from its separate sin and cos by an ulp for some arguments."""
    sn, cs = ctypes.c_double(), ctypes.c_double()
    _libm.sincos(x, ctypes.byref(sn), ctypes.byref(cs))
    return sn.value, cs.value


def rem_euclid(x: float, m: float) -> float:
    """Rust's f64::rem_euclid: the remainder of x / m in [0, |m|)."""
    r = math.fmod(x, m)
    return r + abs(m) if r < 0.0 else r


def to_degrees(x: float) -> float:
    return x * (180.0 / math.pi)


def to_radians(x: float) -> float:
    return x * (math.pi / 180.0)

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
    sn, cs = sincos(n_rad)
    f = 1.0004 - 0.0373 * cs + 0.0002 * sincos(2.0 * n_rad)[1]
    u = -2.14 * sn
    return f, u


OBLIQUITY_DEG = 23.452          # astronomy.rs:479-480: the ecliptic's obliquity and the Moon's
LUNAR_INCLINATION_DEG = 5.145   # orbital inclination (Schureman 1958)


def lunar_intersection(n_rad: float):
    """The Moon's orbital inclination I to the equator and the longitude xi of
    its intersection with the equator (radians) at node longitude N, Schureman
    (1958) eqs. 191-197 (astronomy.rs:488-500)."""
    w, i = to_radians(OBLIQUITY_DEG), to_radians(LUNAR_INCLINATION_DEG)
    incl = math.acos(math.cos(w) * math.cos(i) - math.sin(w) * math.sin(i) * math.cos(n_rad))
    n = math.pi - rem_euclid(math.pi - n_rad, 2.0 * math.pi)
    t = math.tan(0.5 * n)
    a = math.atan(math.cos(0.5 * (w - i)) / math.cos(0.5 * (w + i)) * t)
    b = math.atan(math.sin(0.5 * (w - i)) / math.sin(0.5 * (w + i)) * t)
    return incl, n - (a + b)


def nodal_l2_relative(n_rad: float, p_rad: float):
    """L2's factor relative to M2's, 1 - 6 tan^2(I/2) e^{2iP} with P = p - xi:
    (modulus, argument in degrees) (Schureman eqs. 213-214; astronomy.rs:509-516)."""
    incl, xi = lunar_intersection(n_rad)
    tq = math.tan(0.5 * incl)
    k = 6.0 * (tq * tq)
    two_p = 2.0 * (p_rad - xi)
    sn, cs = sincos(two_p)
    re, im = 1.0 - k * cs, -k * sn
    return hypot(re, im), to_degrees(math.atan2(im, re))


def _powi(x: float, k: int) -> float:
    r = x
    for _ in range(k - 1):
        r = r * x
    return r


def nodal_factors(kind, n_rad: float, p_rad: float):
    """f and u (degrees) of a nodal family (astronomy.rs:520-566)."""
    (s, c), (s2, c2), (s3, c3) = sincos(n_rad), sincos(2.0 * n_rad), sincos(3.0 * n_rad)
    if kind is None:
        return 1.0, 0.0
    if kind == "M2":
        return nodal_m2(n_rad)
    if kind == "L2":
        fm, um = nodal_m2(n_rad)
        fr, ur = nodal_l2_relative(n_rad, p_rad)
        return fm * fr, um + ur
    if kind == "K1":
        return 1.0060 + 0.1150 * c - 0.0088 * c2 + 0.0006 * c3, -8.86 * s + 0.68 * s2 - 0.07 * s3
    if kind == "O1":
        return 1.0089 + 0.1871 * c - 0.0147 * c2 + 0.0014 * c3, 10.80 * s - 1.34 * s2 + 0.19 * s3
    if kind == "K2":
        return 1.0241 + 0.2863 * c + 0.0083 * c2 - 0.0015 * c3, -17.74 * s + 0.68 * s2 - 0.04 * s3
    if kind == "Mf":
        return 1.0429 + 0.4135 * c - 0.004 * c2, -23.7 * s + 2.7 * s2 - 0.4 * s3
    if kind == "Mm":
        return 1.0 - 0.1300 * c + 0.0013 * c2, 0.0
    # ("M2^", k): f = f(M2)^k, u = k u(M2)
    fm, um = nodal_m2(n_rad)
    return _powi(fm, kind[1]), float(kind[1]) * um


# name -> (Doodson argument, nodal family) (astronomy.rs:343-458; the 180 degree offsets of
# lambda2 and L2 are Foreman's, the +-90 of the diurnals Schureman's)
CONSTITUENTS = {
    "M2": (Doodson(2.0, 0.0, 0.0, 0.0, 0.0, 0.0), "M2"),
    "S2": (Doodson(2.0, 2.0, -2.0, 0.0, 0.0, 0.0), None),
    "N2": (Doodson(2.0, -1.0, 0.0, 1.0, 0.0, 0.0), "M2"),
    "K2": (Doodson(2.0, 2.0, 0.0, 0.0, 0.0, 0.0), "K2"),
    "2N2": (Doodson(2.0, -2.0, 0.0, 2.0, 0.0, 0.0), "M2"),
    "MU2": (Doodson(2.0, -2.0, 2.0, 0.0, 0.0, 0.0), "M2"),
    "NU2": (Doodson(2.0, -1.0, 2.0, -1.0, 0.0, 0.0), "M2"),
    "LDA2": (Doodson(2.0, 1.0, -2.0, 1.0, 0.0, 180.0), "M2"),
    "L2": (Doodson(2.0, 1.0, 0.0, -1.0, 0.0, 180.0), "L2"),
    "T2": (Doodson(2.0, 2.0, -3.0, 0.0, 1.0, 0.0), None),
    "K1": (Doodson(1.0, 1.0, 0.0, 0.0, 0.0, 90.0), "K1"),
    "O1": (Doodson(1.0, -1.0, 0.0, 0.0, 0.0, -90.0), "O1"),
    "P1": (Doodson(1.0, 1.0, -2.0, 0.0, 0.0, -90.0), None),
    "Q1": (Doodson(1.0, -2.0, 0.0, 1.0, 0.0, -90.0), "O1"),
    "M4": (Doodson(4.0, 0.0, 0.0, 0.0, 0.0, 0.0), ("M2^", 2)),
    "MS4": (Doodson(4.0, 2.0, -2.0, 0.0, 0.0, 0.0), "M2"),
    "MN4": (Doodson(4.0, -1.0, 0.0, 1.0, 0.0, 0.0), ("M2^", 2)),
    "M6": (Doodson(6.0, 0.0, 0.0, 0.0, 0.0, 0.0), ("M2^", 3)),
    "MF": (Doodson(0.0, 2.0, 0.0, 0.0, 0.0, 0.0), "Mf"),
    "MM": (Doodson(0.0, 1.0, 0.0, -1.0, 0.0, 0.0), "Mm"),
    "SSA": (Doodson(0.0, 0.0, 2.0, 0.0, 0.0, 0.0), None),
}
CONSTITUENT_NAMES = ("M2", "S2", "N2", "K2", "2N2", "MU2", "NU2", "LDA2", "L2", "T2", "K1", "O1", "P1", "Q1", "M4",
                     "MS4", "MN4", "M6", "Mf", "Mm", "Ssa")


def canonical_name(name: str):
    """The conventional spelling of a supported constituent (case-insensitive), or None."""
    key = name.strip().upper()
    for c in CONSTITUENT_NAMES:
        if c.upper() == key:
            return c
    return None


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
    f, u = nodal_factors(kind, a.n * DEG_TO_RAD, a.p * DEG_TO_RAD)
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

    def u_rate_rad(self) -> float:
        """The nodal phase's drift in rad/s, which adds to the angular speed (astronomy.rs:264-266)."""
        return self.u_rate_deg * DEG_TO_RAD

    def relative_f_rate(self) -> float:
        """f-dot / f0 per second (astronomy.rs:270-272)."""
        return self.f_rate / self.at_epoch.f


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

    def nodal_correction(self, name: str, t_mid: float) -> NodalCorrection:
        """f and u at t_mid, V0 at the epoch: the correction held over a record (clock.rs:93-101)."""
        at_epoch = nodal_correction(name, self.astronomical_arguments(0.0))
        at_mid = nodal_correction(name, self.astronomical_arguments(t_mid))
        return NodalCorrection(at_mid.f, at_mid.u_deg, at_epoch.v0_deg)

    def nodal_drift(self, name: str) -> NodalDrift:
        """Central differences over +-1 day about the epoch (clock.rs:108-117)."""
        before = nodal_correction(name, self.astronomical_arguments(-DAY))
        at_epoch = nodal_correction(name, self.astronomical_arguments(0.0))
        after = nodal_correction(name, self.astronomical_arguments(DAY))
        return NodalDrift(at_epoch, (after.f - before.f) / (2.0 * DAY), (after.u_deg - before.u_deg) / (2.0 * DAY))


# ---------------------------------------------------------------------------
# Minor constituents by linear admittance (src/tides/admittance.rs)
# ---------------------------------------------------------------------------
# Cartwright-Tayler-Edden equilibrium amplitudes (admittance.rs:40-58)
EQUILIBRIUM_AMPLITUDE = {
    "M2": 0.242_334, "S2": 0.112_841, "N2": 0.046_398, "K2": 0.030_704, "2N2": 0.006_141, "MU2": 0.007_408,
    "NU2": 0.008_811, "LDA2": 0.001_790, "L2": 0.006_863, "T2": 0.006_614, "K1": 0.141_565, "O1": 0.100_514,
    "P1": 0.046_843, "Q1": 0.019_256,
}


def equilibrium_amplitude(name: str):
    return EQUILIBRIUM_AMPLITUDE.get(name.strip().upper())


@dataclass(frozen=True)
class AdmittanceInference:
    """Constituent `name` from `lower` and `upper` by an admittance linear in
    frequency (admittance.rs:63-137)."""
    name: str
    lower: str
    upper: str

    def weights(self):
        """(w_lower, w_upper) of Z = w_lower Z_lower + w_upper Z_upper, or None."""
        try:
            om, oa, ob = angular_speed(self.name), angular_speed(self.lower), angular_speed(self.upper)
        except KeyError:
            return None
        if oa == ob:
            return None
        e, ea, eb = (equilibrium_amplitude(n) for n in (self.name, self.lower, self.upper))
        if e is None or ea is None or eb is None:
            return None
        x = (om - oa) / (ob - oa)
        return (1.0 - x) * e / ea, x * e / eb

    def infer(self, lower, upper):
        """The inferred (H, G) (G in degrees in [0, 360)) from the references' (H, G)."""
        w = self.weights()
        if w is None:
            return None
        wa, wb = w

        def complex_(hg):
            h, g = hg
            sn, cs = sincos(to_radians(g))
            return h * cs, -h * sn

        a, b = complex_(lower), complex_(upper)
        re, im = wa * a[0] + wb * b[0], wa * a[1] + wb * b[1]
        return hypot(re, im), rem_euclid(to_degrees(math.atan2(-im, re)), 360.0)


MINOR_SEMIDIURNALS = (AdmittanceInference("2N2", "N2", "M2"), AdmittanceInference("MU2", "N2", "M2"),
                      AdmittanceInference("NU2", "N2", "M2"), AdmittanceInference("LDA2", "M2", "S2"),
                      AdmittanceInference("L2", "M2", "S2"))
MINORS = MINOR_SEMIDIURNALS + (AdmittanceInference("T2", "M2", "S2"),)
