"""Tidal atlases as the open boundary's tides per node (src/boundary/tidal_atlas.rs,
src/boundary/coverage.rs, src/io/projection.rs): the host side of the atlas
forcing, which hands the device its per-node constants (`cases.BoundaryTides`,
the form `dg_set_boundary_tides` uploads).

An atlas holds, at geographic points, each constituent's amplitude and Greenwich
lag of the elevation and (optionally) of the depth-mean eastward and northward
velocity, plus the source model's depth. Projected onto a placed mesh's open
boundary nodes: the complex amplitudes H e^{-iG} are interpolated with
inverse-distance weights (power 2) from the 4 nearest points, the velocity turned
from east/north into the mesh axes, and each constituent's nodal terms at the
run's epoch folded in, so that a node's tide is
R(t) sum_j f_j(t)/f0_j (C_j cos w'_j t + S_j sin w'_j t) for eta, u and v.
The arithmetic is the reference's, operation for operation, in IEEE doubles with
the platform libm's trigonometry (which the reference calls too).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import cases, tides as T

NEIGHBOURS = 4                  # tidal_atlas.rs:67
EARTH_RADIUS = 6_371_000.0      # tidal_atlas.rs:518 (TidalAtlas::nearest)


class TidalAtlasError(Exception):
    """An atlas that cannot be read or used (the reference's TidalAtlasError)."""


# ---------------------------------------------------------------------------
# The local projection (LocalProjection, src/io/projection.rs:97-172)
# ---------------------------------------------------------------------------
class LocalProjection:
    """Equirectangular about a reference point, with WGS84's radii of curvature there."""
    A = 6_378_137.0
    F = 1.0 / 298.257_223_563

    def __init__(self, ref_lat: float, ref_lon: float):
        lat_rad = ref_lat * math.pi / 180.0
        sin_lat, cos_lat = T.sincos(lat_rad)
        e2 = 2.0 * self.F - self.F * self.F
        sin2 = sin_lat * sin_lat
        rho = self.A * (1.0 - e2) / math.pow(1.0 - e2 * sin2, 1.5)
        nu = self.A / math.sqrt(1.0 - e2 * sin2)
        self.ref_lat, self.ref_lon = float(ref_lat), float(ref_lon)
        self.meters_per_deg_lat = rho * math.pi / 180.0
        self.meters_per_deg_lon = nu * cos_lat * math.pi / 180.0

    def geo_to_xy(self, lat: float, lon: float):
        return (lon - self.ref_lon) * self.meters_per_deg_lon, (lat - self.ref_lat) * self.meters_per_deg_lat

    def xy_to_geo(self, x: float, y: float):
        return self.ref_lat + y / self.meters_per_deg_lat, self.ref_lon + x / self.meters_per_deg_lon


# ---------------------------------------------------------------------------
# The atlas (TidalAtlas, tidal_atlas.rs:112-560)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class AtlasConstituent:
    name: str
    eta: tuple                  # (amplitude m, lag degrees)
    velocity: tuple | None      # ((east amp, lag), (north amp, lag)) in m/s and degrees


@dataclasses.dataclass
class AtlasPoint:
    lon: float
    lat: float
    depth: float
    mean: float | None
    constituents: list


@dataclasses.dataclass
class TidalAtlas:
    points: list = dataclasses.field(default_factory=list)
    header: list = dataclasses.field(default_factory=list)

    @staticmethod
    def read(path: str) -> "TidalAtlas":
        with open(path) as f:
            return TidalAtlas.parse(f.read())

    @staticmethod
    def parse(text: str) -> "TidalAtlas":
        """Whitespace-separated `lon lat depth name eta_amp eta_lag [u_amp u_lag v_amp v_lag]`,
        `#` comments, a point's lines consecutive, `Z0` its mean level (tidal_atlas.rs:197-262)."""
        atlas = TidalAtlas()
        for i, raw in enumerate(text.splitlines()):
            line = raw.strip()
            if line.startswith("#"):
                if not atlas.points:
                    atlas.header.append(line[1:].strip())
                continue
            if not line:
                continue
            fields = line.split()
            if len(fields) not in (6, 10):
                raise TidalAtlasError(f"line {i + 1}: expected 6 or 10 columns, found {len(fields)}")

            def num(k):
                try:
                    return float(fields[k])
                except ValueError:
                    raise TidalAtlasError(f"line {i + 1}: bad number {fields[k]!r}") from None

            lon, lat, depth = num(0), num(1), num(2)
            if not atlas.points or atlas.points[-1].lon != lon or atlas.points[-1].lat != lat:
                atlas.points.append(AtlasPoint(lon, lat, depth, None, []))
            point = atlas.points[-1]
            if fields[3].upper() == "Z0":
                point.mean = num(4)
                continue
            name = T.canonical_name(fields[3])
            if name is None:
                raise TidalAtlasError(f"line {i + 1}: unsupported constituent {fields[3]}")
            velocity = ((num(6), num(7)), (num(8), num(9))) if len(fields) == 10 else None
            point.constituents.append(AtlasConstituent(name, (num(4), num(5)), velocity))
        atlas.validate()
        return atlas

    def validate(self):
        """Points, and the same constituents at every one, with velocity everywhere or nowhere."""
        if not self.points:
            raise TidalAtlasError("no points")

        def signature(p):
            return [(c.name, c.velocity is not None) for c in p.constituents]

        expected = signature(self.points[0])
        if not expected:
            raise TidalAtlasError("no constituents")
        if any(v != expected[0][1] for _, v in expected):
            raise TidalAtlasError("velocity given for some constituents only")
        for p in self.points:
            if signature(p) != expected:
                raise TidalAtlasError(f"point ({p.lon}, {p.lat}) has constituents {signature(p)}, expected {expected}")

    def names(self) -> list:
        return [c.name for c in self.points[0].constituents] if self.points else []

    def has_velocity(self) -> bool:
        return bool(self.points and self.points[0].constituents and self.points[0].constituents[0].velocity)

    def infer_by_admittance(self, inference: T.AdmittanceInference):
        """Replace or add `inference.name` at every point from its two references, elevation
        and velocity alike (tidal_atlas.rs:352-398)."""
        name = T.canonical_name(inference.name)
        if name is None:
            raise TidalAtlasError(f"unknown constituent {inference.name}")
        if inference.weights() is None:
            raise TidalAtlasError(f"no admittance inference of {name} from {inference.lower} and {inference.upper}")
        for p in self.points:
            def find(source):
                for c in p.constituents:
                    if c.name.upper() == source.upper():
                        return c
                raise TidalAtlasError(f"{name} is inferred from {source}, which is missing")

            a, b = find(inference.lower), find(inference.upper)
            vel = None
            if a.velocity is not None and b.velocity is not None:
                vel = (inference.infer(a.velocity[0], b.velocity[0]), inference.infer(a.velocity[1], b.velocity[1]))
            inferred = AtlasConstituent(name, inference.infer(a.eta, b.eta), vel)
            for q, c in enumerate(p.constituents):
                if c.name == name:
                    p.constituents[q] = inferred
                    break
            else:
                p.constituents.append(inferred)

    def infer(self, name: str, source: str, ratio: float, lag_offset_deg: float):
        """Replace or add `name` at every point by `source` scaled by `ratio` and lagged by
        `lag_offset_deg`, elevation and velocity alike (tidal_atlas.rs:315-347)."""
        canon = T.canonical_name(name)
        if canon is None:
            raise TidalAtlasError(f"unknown constituent {name}")

        def scaled(hg):
            return ratio * hg[0], T.rem_euclid(hg[1] + lag_offset_deg, 360.0)

        for p in self.points:
            src = next((c for c in p.constituents if c.name.upper() == source.upper()), None)
            if src is None:
                raise TidalAtlasError(f"{canon} is inferred from {source}, which is missing")
            vel = None if src.velocity is None else (scaled(src.velocity[0]), scaled(src.velocity[1]))
            inferred = AtlasConstituent(canon, scaled(src.eta), vel)
            for q, c in enumerate(p.constituents):
                if c.name == canon:
                    p.constituents[q] = inferred
                    break
            else:
                p.constituents.append(inferred)

    def infer_from_theory(self):
        """K2 and P1 at their equilibrium ratios to S2 and K1 with equal lags, and the minor
        semidiurnals by admittance (tidal_atlas.rs:410-422)."""
        for name, source in (("K2", "S2"), ("P1", "K1")):
            self.infer(name, source, T.equilibrium_amplitude(name) / T.equilibrium_amplitude(source), 0.0)
        for inference in T.MINOR_SEMIDIURNALS:
            self.infer_by_admittance(inference)

    def infer_missing_minors(self) -> list:
        """2N2, mu2, nu2, lambda2, L2 and T2 where the atlas lacks them, by admittance from
        its M2, S2 and N2; what it carries is kept (tidal_atlas.rs:429-445)."""
        present = [n.upper() for n in self.names()]
        added = []
        for inference in T.MINORS:
            if inference.name.upper() in present:
                continue
            self.infer_by_admittance(inference)
            added.append(T.canonical_name(inference.name))
        return added

    def nearest(self, lon: float, lat: float):
        """The nearest point and its distance (m, spherical earth) (tidal_atlas.rs:518-530)."""
        best = None
        for p in self.points:
            (s1, c1), (s2, c2) = T.sincos(T.to_radians(lat)), T.sincos(T.to_radians(p.lat))
            dlon = T.to_radians(p.lon - lon)
            c = s1 * s2 + c1 * c2 * math.cos(dlon)
            d = EARTH_RADIUS * math.acos(min(max(c, -1.0), 1.0))
            if best is None or d < best[1]:
                best = (p, d)
        return best

    def tides_at(self, positions, projection: LocalProjection, clock: T.ModelClock, coverage_radius: float):
        """Per position (mesh coordinates): the constants C, S of eta, u, v per constituent at the
        epoch's nodal terms [n_c][6], the mean level and the source depth; per constituent
        w' = w + u-dot and f-dot / f0 (tidal_atlas.rs:637-776)."""
        self.validate()
        names = self.names()
        velocity = self.has_velocity()
        drifts = [clock.nodal_drift(n) for n in names]
        omega = [2.0 * math.pi / T.constituent_period(n) + d.u_rate_rad() for n, d in zip(names, drifts)]
        f_rate = [d.relative_f_rate() for d in drifts]
        xy = [projection.geo_to_xy(p.lat, p.lon) for p in self.points]
        coefficients = np.zeros((len(positions), len(names), 6))
        mean = np.zeros(len(positions))
        source_depth = np.zeros(len(positions))
        for s, (x, y) in enumerate(positions):
            dist = sorted(((T.hypot(px - x, py - y), i) for i, (px, py) in enumerate(xy)), key=lambda a: a[0])
            nearest = dist[:NEIGHBOURS]
            d_min = nearest[0][0]
            if d_min > coverage_radius:
                lat, lon = projection.xy_to_geo(x, y)
                raise TidalAtlasError(f"the boundary node at ({x}, {y}) (lat {lat}, lon {lon}) is {d_min} m from the "
                                      f"atlas, beyond its coverage radius {coverage_radius} m")
            if d_min < 1.0:
                weights = [(1.0, nearest[0][1])]
            else:
                w = [(1.0 / (d * d), i) for d, i in nearest]
                total = 0.0
                for wi, _ in w:
                    total += wi
                weights = [(wi / total, i) for wi, i in w]
            # the direction of local north in the projected plane
            lat, lon = projection.xy_to_geo(x, y)
            nx, ny = projection.geo_to_xy(lat + 1e-4, lon)
            theta = math.atan2(nx - x, ny - y)
            sin_t, cos_t = T.sincos(theta)
            acc = 0.0
            for w, i in weights:
                acc += w * (self.points[i].mean if self.points[i].mean is not None else 0.0)
            mean[s] = acc
            acc = 0.0
            for w, i in weights:
                acc += w * self.points[i].depth
            source_depth[s] = acc
            for j, drift in enumerate(drifts):
                n = drift.at_epoch

                def interp(get):
                    re, im = 0.0, 0.0
                    for w, i in weights:
                        amp, lag = get(self.points[i].constituents[j])
                        sg, cg = T.sincos(T.to_radians(lag))
                        re, im = re + w * amp * cg, im - w * amp * sg
                    return re, im

                eta = interp(lambda c: c.eta)
                if velocity:
                    east = interp(lambda c: c.velocity[0])
                    north = interp(lambda c: c.velocity[1])
                    u = (east[0] * cos_t + north[0] * sin_t, east[1] * cos_t + north[1] * sin_t)
                    v = (-east[0] * sin_t + north[0] * cos_t, -east[1] * sin_t + north[1] * cos_t)
                else:
                    u = v = (0.0, 0.0)
                sin_p, cos_p = T.sincos(n.phase_offset_rad())
                for q, (re, im) in enumerate((eta, u, v)):
                    re, im = re * cos_p - im * sin_p, re * sin_p + im * cos_p
                    coefficients[s, j, 2 * q] = n.f * re
                    coefficients[s, j, 2 * q + 1] = -n.f * im
        return np.array(omega), np.array(f_rate), coefficients, mean, source_depth, velocity


# ---------------------------------------------------------------------------
# Coverage of the open boundary (BoundaryCoverage::of_atlas, coverage.rs:73-151,191-207)
# ---------------------------------------------------------------------------
def uncovered_faces(case, projection: LocalProjection, atlas: TidalAtlas, max_distance: float, tag=cases.TIDE):
    """The open faces with a node farther than `max_distance` from every atlas point, and the
    lengths of the tagged faces and of those."""
    n1 = case.order + 1
    x, y = cases.node_coordinates(case.mesh, case.order)
    x, y = x.reshape(case.n_elements, -1), y.reshape(case.n_elements, -1)
    c = case.mesh.corners
    out, length, missed_length = [], 0.0, 0.0
    for k in range(case.n_elements):
        for f in range(4):
            if case.mesh.links[k, f] != cases.BOUNDARY or case.mesh.tags[k, f] != tag:
                continue
            x0, y0 = c[k, f]
            x1_, y1_ = c[k, (f + 1) % 4]
            face_length = T.hypot(float(x1_ - x0), float(y1_ - y0))
            length += face_length
            missed = False
            for fi in range(n1):
                nd = cases.face_node(n1, f, fi)
                lat, lon = projection.xy_to_geo(float(x[k, nd]), float(y[k, nd]))
                near = atlas.nearest(lon, lat)
                if near is None or not near[1] <= max_distance:
                    missed = True
            if missed:
                out.append((k, f))
                missed_length += face_length
    return out, length, missed_length


def close_uncovered(case, faces, length, missed_length, max_fraction: float):
    """Make the uncovered faces walls, refused when they exceed `max_fraction` of the boundary's
    length (BoundaryCoverage::close)."""
    share = missed_length / length if length > 0.0 else 0.0
    if share > max_fraction:
        raise TidalAtlasError(f"{len(faces)} open faces ({missed_length:.0f} of {length:.0f} m) are not covered "
                              f"by the atlas, more than the share {max_fraction} that may become walls")
    for k, f in faces:
        case.mesh.tags[k, f] = cases.WALL
    return len(faces)


def boundary_tides(case, atlas: TidalAtlas, projection: LocalProjection, clock: T.ModelClock, coverage_radius: float,
                   ramp: float = 0.0, max_ratio: float = 0.0, use_mean: bool = False) -> cases.BoundaryTides:
    """The atlas at every open-boundary face node of `case` (TidalAtlas::boundary_tides with
    with_ramp_up and with_transport_scaling): one slot per face node, positions by the
    element's bilinear map."""
    nodes = cases.open_face_nodes(case)
    omega, f_rate, coef, mean, depth, velocity = atlas.tides_at([(x, y) for _, _, _, x, y in nodes], projection,
                                                                clock, coverage_radius)
    n1 = case.order + 1
    slot = np.full((case.n_elements, 4, n1), 0xFFFFFFFF, dtype=np.uint32)
    for i, (k, f, fi, _, _) in enumerate(nodes):
        slot[k, f, fi] = i
    return cases.BoundaryTides(omega, f_rate, slot, coef, mean if use_mean else np.zeros_like(mean), depth, ramp,
                               velocity, max_ratio)


def evaluate(bt: cases.BoundaryTides, slot: int, t: float):
    """(eta, u, v) of slot `slot` at time t, the nodal terms drifted to t and the ramp applied
    (BoundaryTides::evaluate, tidal_atlas.rs:948-963; a host check of the constants)."""
    s0 = s1 = s2 = 0.0
    for j in range(bt.omega.size):
        f = 1.0 + bt.f_rate[j] * t
        cs, sn = math.cos(bt.omega[j] * t), math.sin(bt.omega[j] * t)
        c = bt.coef[slot, j]
        s0 += f * (c[0] * cs + c[1] * sn)
        s1 += f * (c[2] * cs + c[3] * sn)
        s2 += f * (c[4] * cs + c[5] * sn)
    if bt.ramp > 0.0:
        tau = min(max(t / bt.ramp, 0.0), 1.0)
        ramp = tau * tau * (3.0 - 2.0 * tau)
    else:
        ramp = 1.0
    return bt.mean[slot] + ramp * s0, ramp * s1, ramp * s2
