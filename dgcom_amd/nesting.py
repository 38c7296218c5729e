"""A parent ocean model sampled at a mesh's open boundary and relaxation band
(src/io/geo_grid.rs, src/io/ocean_model.rs, src/io/field_series.rs,
src/boundary/band.rs, src/boundary/ocean_nesting.rs:174-420,756-800): the host
side of nesting, which hands the device its per-node form (`cases.Nesting`, the
tables `dg_set_nesting` uploads).

The parent's fields (f32 snapshots of the surface, the east and north current,
and its depth) live on a regular longitude/latitude grid. Every mesh node on an
open face, and every node within the band's width of one (the distance to the
open faces as polylines through their nodes), takes the bilinear stencil of the
parent cell holding it, renormalised over the cell's wet corners (or the nearest
wet point within `max_snap` for a boundary node on parent land). Its series of
(zeta, u, v) per snapshot is interpolated in space once and kept in f32, the
current turned from east/north into the mesh axes; its band weight is 1 on the
boundary and the cosine profile of 1 - d / width inside. The bed can be blended
towards the parent's depth across the band, shores kept. The arithmetic is the
reference's operation for operation, with the platform libm (which it calls too).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import cases, tides as T

EARTH_RADIUS = 6_371_000.0


class NestingError(Exception):
    """A parent that cannot serve the mesh (the reference's NestingError)."""


# ---------------------------------------------------------------------------
# The parent's grid (GeoGrid, geo_grid.rs)
# ---------------------------------------------------------------------------
def _bracket(coords, value):
    """The cell index along one axis and the local coordinate in it (geo_grid.rs:482-502)."""
    n = len(coords)
    ascending = coords[1] > coords[0]
    lo, hi = (coords[0], coords[n - 1]) if ascending else (coords[n - 1], coords[0])
    if not (lo <= value <= hi):
        return None
    if ascending:
        past = next((i for i, c in enumerate(coords) if not c <= value), n)
    else:
        past = next((i for i, c in enumerate(coords) if not c >= value), n)
    k = min(max(past, 1), n - 1) - 1
    f = (value - coords[k]) / (coords[k + 1] - coords[k])
    return k, min(max(f, 0.0), 1.0)


def geodesic_distance(a, b) -> float:
    """Great-circle distance (m) between (lon, lat) points (geo_grid.rs)."""
    (lon1, lat1), (lon2, lat2) = a, b
    p1, p2 = T.to_radians(lat1), T.to_radians(lat2)
    dp, dl = p2 - p1, T.to_radians(lon2 - lon1)
    s1, s2 = math.sin(0.5 * dp), math.sin(0.5 * dl)
    h = s1 * s1 + math.cos(p1) * math.cos(p2) * (s2 * s2)
    return 2.0 * EARTH_RADIUS * math.asin(min(math.sqrt(h), 1.0))


@dataclasses.dataclass
class GeoGrid:
    """A regular grid of longitudes `lon` (nx) and latitudes `lat` (ny), points row by row."""
    lon: list
    lat: list

    @property
    def nx(self):
        return len(self.lon)

    @property
    def ny(self):
        return len(self.lat)

    def __len__(self):
        return self.nx * self.ny

    def position(self, k: int):
        return self.lon[k % self.nx], self.lat[k // self.nx]

    def locate(self, lon: float, lat: float):
        """(j, i, fx, fy) of the cell holding the point, or None outside the grid."""
        if not (math.isfinite(lon) and math.isfinite(lat)):
            return None
        bi = _bracket(self.lon, lon)
        bj = _bracket(self.lat, lat)
        if bi is None or bj is None:
            return None
        return bj[0], bi[0], bi[1], bj[1]

    def corners(self, j: int, i: int):
        base = j * self.nx + i
        return [base, base + 1, base + self.nx, base + self.nx + 1]

    def nearest(self, lon, lat, max_distance, accept):
        best = None
        for k in range(len(self)):
            if not accept(k):
                continue
            d = geodesic_distance((lon, lat), self.position(k))
            if d <= max_distance and (best is None or d < best[1]):
                best = (k, d)
        return best


def bilinear_stencil(grid: GeoGrid, cell, valid):
    """Bilinear weights renormalised over the valid corners, invalid corners pointing at a
    valid one with weight 0 (Stencil::bilinear, geo_grid.rs:420-450); None with none valid."""
    j, i, fx, fy = cell
    idx = grid.corners(j, i)
    w = [(1.0 - fx) * (1.0 - fy), fx * (1.0 - fy), (1.0 - fx) * fy, fx * fy]
    w = [wk if valid(k) else 0.0 for wk, k in zip(w, idx)]
    total = 0.0
    for wk in w:
        total += wk
    fallback = next((k for k in idx if valid(k)), None)
    if fallback is None:
        return None
    if total <= 1e-12:
        cands = [c for c in range(4) if valid(idx[c])]
        c = min(cands, key=lambda c: T.hypot(fx - (c % 2), fy - (c // 2)))
        return [idx[c]] * 4, [1.0, 0.0, 0.0, 0.0]
    return [k if valid(k) else fallback for k in idx], [wk / total for wk in w]


# ---------------------------------------------------------------------------
# The parent (OceanModelReader, ocean_model.rs:85-170,275-380)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class ParentModel:
    """Snapshots at Unix `times` of `ssh`, `east`, `north` ([n_times][n_points] f32) on `grid`,
    and the parent's still-water `depth` [n_points] (or None)."""
    grid: GeoGrid
    times: np.ndarray
    ssh: np.ndarray
    east: np.ndarray | None = None
    north: np.ndarray | None = None
    depth: np.ndarray | None = None

    def __post_init__(self):
        self.times = np.asarray(self.times, dtype=np.float64)
        if self.times.size == 0 or np.any(np.diff(self.times) <= 0.0) or not np.all(np.isfinite(self.times)):
            raise NestingError("the parent's times must be finite and strictly increasing")
        fields = [f for f in (self.ssh, self.east, self.north) if f is not None]
        self.wet = np.array([(self.depth is None or self.depth[k] > 0.0)
                             and all(np.all(np.isfinite(f[:, k])) for f in fields) for k in range(len(self.grid))])

    def has_velocity(self):
        return self.east is not None and self.north is not None

    def stencil(self, lon, lat):
        cell = self.grid.locate(lon, lat)
        if cell is None:
            return None
        return bilinear_stencil(self.grid, cell, lambda k: bool(self.wet[k]))

    def interpolate(self, field, t, stencil):
        """One snapshot's value at a stencil: sum of w f32-as-f64 (FieldSeries::interpolate)."""
        idx, w = stencil
        value = 0.0
        for k, wk in zip(idx, w):
            value += wk * float(field[t, k])
        return 0.0 + 1.0 * value


# ---------------------------------------------------------------------------
# Distances to the open boundary (band.rs:15-125)
# ---------------------------------------------------------------------------
def point_segment_distance(p, a, b) -> float:
    dx, dy = b[0] - a[0], b[1] - a[1]
    len2 = dx * dx + dy * dy
    t = min(max(((p[0] - a[0]) * dx + (p[1] - a[1]) * dy) / len2, 0.0), 1.0) if len2 > 0.0 else 0.0
    return T.hypot(p[0] - a[0] - t * dx, p[1] - a[1] - t * dy)


def boundary_distances(case, band_width: float, tag=cases.TIDE):
    """Per flat node: 0 on an open face, else its distance to the open faces (polylines through
    their nodes) where within `band_width`, else infinity."""
    n1, nn = case.order + 1, case.nn
    x, y = cases.node_coordinates(case.mesh, case.order)
    x, y = x.ravel(), y.ravel()
    dist = np.full(case.n_elements * nn, np.inf)
    segments = []
    for k in range(case.n_elements):
        for f in range(4):
            if case.mesh.links[k, f] != cases.BOUNDARY or case.mesh.tags[k, f] != tag:
                continue
            nodes = [k * nn + cases.face_node(n1, f, fi) for fi in range(n1)]
            for a, b in zip(nodes[:-1], nodes[1:]):
                segments.append(((float(x[a]), float(y[a])), (float(x[b]), float(y[b]))))
            dist[nodes] = 0.0
    if not segments:
        raise NestingError("no open boundary to nest")
    if band_width > 0.0:
        for n in range(dist.size):
            if dist[n] > 0.0:
                p = (float(x[n]), float(y[n]))
                best = math.inf
                for a, b in segments:
                    best = min(best, point_segment_distance(p, a, b))
                dist[n] = best if best <= band_width else math.inf
    return dist


def east_axis(projection, lat, lon):
    """East in the mesh plane, (cos theta, sin theta) (geo_grid.rs:374-384)."""
    x0, y0 = projection.geo_to_xy(lat, lon)
    x1, y1 = projection.geo_to_xy(lat + 1e-4, lon)
    nx, ny = x1 - x0, y1 - y0
    n = T.hypot(nx, ny)
    return ny / n, -nx / n


# ---------------------------------------------------------------------------
# The parent at the mesh (OceanModelState::new, ocean_nesting.rs:312-420)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class ParentAtMesh:
    """The nesting form and what the bed's blending needs: per slot its position, whether it is
    on the open boundary, and the elements with a relaxed node."""
    nesting: cases.Nesting
    position: list
    on_boundary: list
    band_elements: np.ndarray
    snapped: int


def parent_at_mesh(parent: ParentModel, case, projection, epoch: float, band_width: float = 0.0,
                   profile=cases.cosine_profile, max_snap: float = 2000.0, cubic: bool = True,
                   transport_scaling: bool = True, max_ratio: float = 3.0, reference_level: float = 0.0,
                   ramp: float = 0.0, timescale: float = 3600.0, momentum_only: bool = False,
                   wet_depth: float = 0.05) -> ParentAtMesh:
    """Sample `parent` at the open boundary's nodes and the band's (OceanModelState::new with
    a ParentDatum::Shift datum `reference_level`), with the relaxation's timescale, wet depth
    and momentum switch (OceanModelState::relaxation)."""
    nn = case.nn
    x, y = cases.node_coordinates(case.mesh, case.order)
    x, y = x.ravel(), y.ravel()
    dist = boundary_distances(case, band_width)
    slot = np.full(dist.size, 0xFFFFFFFF, dtype=np.uint32)
    band = np.zeros(case.n_elements, dtype=bool)
    stencils, easts, depths, weights, positions, on_b = [], [], [], [], [], []
    snapped = 0
    for n, d in enumerate(dist):
        on_boundary = d == 0.0
        if not (on_boundary or d < band_width):
            continue
        px, py = float(x[n]), float(y[n])
        lat, lon = projection.xy_to_geo(px, py)
        st = parent.stencil(lon, lat)
        if st is None:
            near = parent.grid.nearest(lon, lat, max_snap, lambda k: bool(parent.wet[k]))
            if near is not None:
                snapped += int(on_boundary)
                st = ([near[0]] * 4, [1.0, 0.0, 0.0, 0.0])
        if st is None:
            if on_boundary:
                raise NestingError(f"the open boundary node at ({px}, {py}) (lat {lat}, lon {lon}) is not in a "
                                   f"parent cell with a wet corner nor within {max_snap} m of a wet parent point")
            continue
        if parent.depth is None:
            pdepth = math.nan
        else:
            pdepth = 0.0
            for k, w in zip(*st):
                pdepth += w * float(parent.depth[k])
        weight = 1.0 if on_boundary else profile(1.0 - d / band_width)
        slot[n] = len(weights)
        if weight > 0.0:
            band[n // nn] = True
        stencils.append(st)
        easts.append(east_axis(projection, lat, lon))
        depths.append(pdepth)
        weights.append(weight)
        positions.append((px, py))
        on_b.append(on_boundary)
    nt = parent.times.size
    series = np.zeros((len(weights), nt, 3), dtype=np.float32)
    velocity = parent.has_velocity()
    for s, (st, (c, sn)) in enumerate(zip(stencils, easts)):
        for t in range(nt):
            zeta = parent.interpolate(parent.ssh, t, st)
            if velocity:
                e = parent.interpolate(parent.east, t, st)
                no = parent.interpolate(parent.north, t, st)
                u, v = e * c - no * sn, e * sn + no * c
            else:
                u = v = 0.0
            series[s, t] = (zeta, u, v)
    ns = cases.Nesting(slot, parent.times.copy(), series, np.asarray(weights), np.asarray(depths),
                       timescale=timescale, epoch=epoch, cubic=cubic, wet_depth=wet_depth,
                       momentum_only=momentum_only, velocity=velocity,
                       max_ratio=max_ratio if transport_scaling else 0.0, reference_level=reference_level, ramp=ramp)
    return ParentAtMesh(ns, positions, on_b, band, snapped)


def blend_bathymetry(pm: ParentAtMesh, bed: np.ndarray, nn: int, shore_depth: float) -> int:
    """The bed blended towards the parent's depth across the band, b <- (1 - w) b - w D_parent,
    except at nodes (by position to the millimetre) of an element with a node shallower than
    `shore_depth` (OceanModelState::blend_bathymetry, ocean_nesting.rs:756-798); in place, the
    count of nodes changed."""
    def at(p):
        return (int(round_half_away(p[0] * 1e3)), int(round_half_away(p[1] * 1e3)))

    ns = pm.nesting
    shore = set()
    for n in np.nonzero(ns.slot != 0xFFFFFFFF)[0]:
        k = n // nn
        if np.any(bed[k * nn:(k + 1) * nn] > -shore_depth):
            shore.add(at(pm.position[ns.slot[n]]))
    changed = 0
    for n in np.nonzero(ns.slot != 0xFFFFFFFF)[0]:
        s = ns.slot[n]
        w, dp, b = float(ns.weight[s]), float(ns.depth[s]), float(bed[n])
        if w > 0.0 and dp > 0.0 and b < 0.0 and at(pm.position[s]) not in shore:
            bed[n] = (1.0 - w) * b - w * dp
            changed += 1
    return changed


def round_half_away(v: float) -> float:
    """Rust's f64::round: to the nearest integer, halves away from zero."""
    a = abs(v)
    r = math.floor(a)
    if a - r >= 0.5:
        r += 1.0
    return math.copysign(r, v)


# ---------------------------------------------------------------------------
# The reference's nesting gate (gpu/src/nest.rs): the tidal channel at 500 m nested in
# a made-up parent held in memory
# ---------------------------------------------------------------------------
GATE_T0 = 1_750_000_000.0     # the parent's first snapshot (Unix s), the run's time 0
GATE_HOURS = 8.0
GATE_VARIANTS = {
    # cubic in time, transport-scaled (max 3), datum shifted by -0.04 m, ramp 1 h, relaxed at 30 min
    "cubic": dict(cubic=True, transport_scaling=True, reference_level=-0.04, ramp=3600.0, timescale=1800.0,
                  momentum_only=False),
    # linear, unscaled, no shift, ramp 30 min, the momentum alone relaxed at 15 min
    "linear": dict(cubic=False, transport_scaling=False, reference_level=0.0, ramp=1800.0, timescale=900.0,
                   momentum_only=True),
}


def gate_parent(projection) -> ParentModel:
    """The made-up parent of nest.rs:48-92: an M2-like wave entering eastward, a
    quarter-diurnal and a mean on a 21 x 16 grid of 800 m around the channel's mouth,
    hourly over the run, its depth 36 m + 0.0006 x."""
    xs = [-6000.0 + 800.0 * i for i in range(21)]
    ys = [-4000.0 + 800.0 * j for j in range(16)]
    lon = [projection.xy_to_geo(x, 0.0)[1] for x in xs]
    lat = [projection.xy_to_geo(0.0, y)[0] for y in ys]
    times = [GATE_T0 + 3600.0 * k for k in range(int(GATE_HOURS) + 2)]
    omega = 2.0 * math.pi / 44_714.0
    kappa = omega / math.sqrt(9.81 * 30.0)
    m = len(xs) * len(ys)
    zeta = np.zeros((len(times), m), dtype=np.float32)
    east = np.zeros((len(times), m), dtype=np.float32)
    north = np.zeros((len(times), m), dtype=np.float32)
    for it, tu in enumerate(times):
        t = tu - GATE_T0
        q = 0
        for y in ys:
            for x in xs:
                phase = omega * t - kappa * x - 0.15 * y / 4000.0
                sp, cp = T.sincos(phase)
                zeta[it, q] = 0.9 * cp + 0.2 * math.cos(2.0 * phase + 0.4) + 0.04
                east[it, q] = 0.32 * math.cos(phase + 0.1) + 0.03
                north[it, q] = 0.05 * sp - 0.02 * x / 6000.0
                q += 1
    depth = np.array([36.0 + 0.0006 * x for _ in ys for x in xs])
    return ParentModel(GeoGrid(lon, lat), np.array(times), zeta, east, north, depth)


def gate_case(variant: str):
    """The nested channel of nest.rs:123-170 (`Domain::new`): 80 x 8 P2 elements over 40 x 4 km,
    its west side open, placed at 63.5 N 8 E, the bed from 30 m down to 2 m above mean sea level
    blended to the parent's across a 3 km band (shores of 2 m kept), still water, the physics as
    the case runner builds a nested case (wet/dry, Coriolis at 63.5 N, Manning 0.025, SSP-RK(4,3)
    at CFL 0.5). Returns the case and the parent at its nodes."""
    from .atlas import LocalProjection
    projection = LocalProjection(63.5, 8.0)
    mesh = cases.structured_mesh(0.0, 40_000.0, 0.0, 4000.0, 80, 8,
                                 sides=(cases.WALL, cases.WALL, cases.WALL, cases.TIDE))
    order = 2
    x, _ = cases.node_coordinates(mesh, order)
    bed = (-(30.0 - 32.0 * x / 40_000.0)).ravel()
    opts = dict(GATE_VARIANTS[variant])
    pm = parent_at_mesh(gate_parent(projection), types_case(mesh, order), projection, GATE_T0, band_width=3000.0,
                        **opts)
    blend_bathymetry(pm, bed, (order + 1) ** 2, 2.0)
    h, hu, hv = cases.still_water(bed)
    case = cases.Case(f"nested channel ({variant})", order, mesh, bed, np.full(mesh.n_elements, 0.025),
                      cases.coriolis_parameter(63.5), cases.Tide(), h, hu, hv, cfl=0.5, scheme=1, n_steps=0)
    return case, pm


def types_case(mesh, order):
    """The mesh and order as the samplers read them (before the bed is known)."""
    import types
    return types.SimpleNamespace(mesh=mesh, order=order, n_elements=mesh.n_elements, nn=(order + 1) ** 2)
