"""The reference's case runner as the caller of this path (src/case): a TOML
case file (src/case/schema.rs) with `--set key=value` overrides, read into a
`cases.Case`, run on the GPU with the runner's output landing, its stations
sampled and written to `stations.csv` (src/case/run.rs:1290-1420), and its
state kept in, and resumed from, the reference's `DGRST2D` restart files
(`dgcom_amd.restart`).

The schema's parts that this path runs: `[run]` (name, start, duration_hours),
`[mesh]` (order, drop_land, origin) with `[mesh.rectangle]`, `[bathymetry.ramp]` or
`[bathymetry.flat]`, `[physics]` (coriolis by latitude or f0, friction by
Manning n), `[[boundary]]` with an elevation-only harmonic tide
(`velocity` "at_rest" or "incoming_wave") or a tidal atlas (`[boundary.atlas]`:
its file, coverage, correction "none", "minors" or "theory", transport scaling
and ramp; `dgcom_amd.atlas`), `[numerics]` (integrator, cfl,
dt_max_s, local_time_stepping_levels), `[output]` (interval_minutes and
`[[output.station]]`). Anything else is refused by name, as the reference's
device path refuses physics it does not run (device.rs:79-143).
"""
from __future__ import annotations

import dataclasses
import datetime as _dt
import math
import os

import numpy as np
import tomli

from . import atlas as A, cases, restart, tides as T

H_VELOCITY = 1e-3            # run.rs:51: WetDryConfig::DEFAULT_H_DRY
DEFAULT_OPEN_WATER_DEPTH = 3.0   # schema.rs:1900
DEFAULT_MAX_OFFSET = 3000.0      # schema.rs (max_offset_m)
LINEAR_CFL = {0: (1.6, 1.04, 0.73, 0.56), 1: (2.2, 1.51, 1.15, 0.93)}   # swe_2d.rs:1263-1302
INSIDE_TOLERANCE = 1e-9      # point_locator.rs:39
LANDING_SLACK = 1e-9         # multirate.rs:588

_KNOWN = {
    "run": {"name", "start", "duration_hours"},
    "mesh": {"order", "drop_land", "rectangle", "origin"},
    "mesh.origin": {"lat", "lon"},
    "mesh.rectangle": {"x_m", "y_m", "nx", "ny", "south", "east", "north", "west"},
    "bathymetry": {"ramp", "flat"},
    "bathymetry.ramp": {"from_m", "to_m", "depth_m"},
    "bathymetry.flat": {"depth_m"},
    "physics": {"coriolis", "friction"},
    "boundary": {"tag", "velocity", "tide", "atlas"},
    "boundary.tide": {"ramp_hours", "constituents", "mean_m"},
    "boundary.atlas": {"file", "coverage_m", "max_closed", "correction", "transport_scaling", "ramp_hours"},
    "numerics": {"integrator", "cfl", "dt_max_s", "local_time_stepping_levels"},
    "output": {"interval_minutes", "station"},
    "output.station": {"name", "x_m", "y_m", "open_water_depth_m", "max_offset_m"},
}


class CaseError(Exception):
    """A case this path cannot run, or an invalid one (the reference's CaseError)."""


def _parse_value(text: str):
    try:
        return tomli.loads(f"v = {text}")["v"]
    except tomli.TOMLDecodeError:
        return text


def apply_overrides(doc: dict, overrides) -> dict:
    """`--set key.path=value` (the value as TOML, else a string); an integer
    path component indexes an array of tables (`boundary.0.velocity`)."""
    for item in overrides:
        if "=" not in item:
            raise CaseError(f"--set {item!r}: expected key=value")
        key, value = item.split("=", 1)
        parts = key.strip().split(".")
        node = doc
        for i, p in enumerate(parts[:-1]):
            if isinstance(node, list):
                node = node[int(p)]
            else:
                nxt = parts[i + 1]
                node = node.setdefault(p, [] if nxt.isdigit() else {})
        last = parts[-1]
        if isinstance(node, list):
            node[int(last)] = _parse_value(value.strip())
        else:
            node[last] = _parse_value(value.strip())
    return doc


def _check_keys(section: str, table: dict):
    known = _KNOWN.get(section)
    if known is None:
        raise CaseError(f"[{section}] is not run by this device path")
    unknown = sorted(set(table) - known)
    if unknown:
        raise CaseError(f"[{section}]: {', '.join(unknown)} not run by this device path")


@dataclasses.dataclass
class Station:
    name: str
    x: float
    y: float
    open_water_depth: float = DEFAULT_OPEN_WATER_DEPTH
    max_offset: float = DEFAULT_MAX_OFFSET


@dataclasses.dataclass
class CaseRun:
    """A case as read: the problem (`case`) and how the runner drives it."""
    name: str
    case: cases.Case
    duration_s: float
    interval_s: float
    stations: list
    dt_max: float | None
    lts_levels: int
    start: tuple | None
    doc: dict = dataclasses.field(default_factory=dict)   # the case as read, overrides applied
    boundary_tides: object = None   # the open boundary's tides per node from an atlas (cases.BoundaryTides)
    notes: list = dataclasses.field(default_factory=list)  # what the loader did (faces closed, minors inferred)


def load(path_or_text: str, overrides=()) -> CaseRun:
    """Read a case file (a path, or the TOML text itself) with overrides."""
    if os.path.exists(path_or_text):
        with open(path_or_text, "rb") as f:
            doc = tomli.load(f)
        base = os.path.dirname(os.path.abspath(path_or_text))
    else:
        doc = tomli.loads(path_or_text)
        base = os.getcwd()
    doc = apply_overrides(doc, overrides)
    for section in doc:
        if section not in ("run", "mesh", "bathymetry", "physics", "boundary", "numerics", "output"):
            raise CaseError(f"[{section}] is not run by this device path")
    run = doc.get("run", {})
    _check_keys("run", run)
    mesh_t = doc.get("mesh", {})
    _check_keys("mesh", mesh_t)
    rect = mesh_t.get("rectangle")
    if rect is None:
        raise CaseError("[mesh]: only [mesh.rectangle] meshes are run by this device path")
    _check_keys("mesh.rectangle", rect)
    order = int(mesh_t.get("order", 2))
    if mesh_t.get("drop_land", False):
        raise CaseError("[mesh] drop_land = true is not run by this device path")
    origin = mesh_t.get("origin")
    if origin is not None:
        _check_keys("mesh.origin", origin)
        origin = (float(origin["lat"]), float(origin["lon"]))
    sides = []
    for side in ("south", "east", "north", "west"):
        kind = rect.get(side, "wall")
        if kind not in ("wall", "open"):
            raise CaseError(f"[mesh.rectangle] {side} = {kind!r}: 'wall' or 'open'")
        sides.append(cases.TIDE if kind == "open" else cases.WALL)
    (x0, x1), (y0, y1) = rect["x_m"], rect["y_m"]
    mesh = cases.structured_mesh(float(x0), float(x1), float(y0), float(y1), int(rect["nx"]), int(rect["ny"]),
                                 sides=tuple(sides))
    x, y = cases.node_coordinates(mesh, order)
    bathy = doc.get("bathymetry", {})
    _check_keys("bathymetry", bathy)
    if "ramp" in bathy:
        r = bathy["ramp"]
        _check_keys("bathymetry.ramp", r)
        bed = -cases.ramp_depth(x, y, tuple(r["from_m"]), tuple(r["to_m"]), tuple(r["depth_m"])).ravel()
    elif "flat" in bathy:
        _check_keys("bathymetry.flat", bathy["flat"])
        bed = np.full(x.size, -float(bathy["flat"]["depth_m"]))
    else:
        raise CaseError("[bathymetry]: a ramp or a flat bed")
    phys = doc.get("physics", {})
    _check_keys("physics", phys)
    cor = phys.get("coriolis")
    f0 = 0.0
    if cor is not None:
        f0 = cases.coriolis_parameter(float(cor["latitude"])) if "latitude" in cor else float(cor.get("f0", 0.0))
    fr = phys.get("friction")
    manning = float(fr["manning"]) if fr is not None else 0.0
    start = None
    if "start" in run:
        st = run["start"]
        if isinstance(st, str):
            st = _dt.datetime.fromisoformat(st.replace("Z", "+00:00"))
        st = st.astimezone(_dt.timezone.utc)
        start = (st.year, st.month, st.day, st.hour, st.minute, st.second)
    tide = cases.Tide()
    incoming = False
    atlas_spec = None
    for b in doc.get("boundary", []):
        _check_keys("boundary", b)
        if b.get("tag", "open") != "open":
            raise CaseError(f"[[boundary]] tag {b.get('tag')!r}: only the rectangle's 'open' sides")
        vel = b.get("velocity", "at_rest")
        if vel not in ("at_rest", "incoming_wave"):
            raise CaseError(f"[[boundary]] velocity = {vel!r}: 'at_rest' or 'incoming_wave'")
        incoming = vel == "incoming_wave"
        t = b.get("tide")
        if t is not None and b.get("atlas") is not None:
            raise CaseError("[[boundary]]: a tide or an atlas, not both")
        if t is not None:
            _check_keys("boundary.tide", t)
            cons = [(c["name"], float(c["amplitude_m"]), float(c.get("phase_deg", 0.0))) for c in t["constituents"]]
            tide = cases.harmonic_tide(cons, float(t.get("ramp_hours", 1.0)), start=start,
                                       mean=float(t.get("mean_m", 0.0)))
        if b.get("atlas") is not None:
            atlas_spec = b["atlas"]
            _check_keys("boundary.atlas", atlas_spec)
    num = doc.get("numerics", {})
    _check_keys("numerics", num)
    integ = num.get("integrator", "ssp_rk43")
    if integ not in ("ssp_rk43", "ssp_rk3"):
        raise CaseError(f"[numerics] integrator = {integ!r}")
    scheme = 1 if integ == "ssp_rk43" else 0
    cfl = float(num["cfl"]) if "cfl" in num else 0.9 * LINEAR_CFL[scheme][order - 1]
    h, hu, hv = cases.still_water(bed)
    case = cases.Case(run.get("name", "case"), order, mesh, bed, np.full(mesh.n_elements, manning), f0, tide,
                      h, hu, hv, cfl=cfl, scheme=scheme, n_steps=0, incoming_wave=incoming)
    notes = []
    boundary_tides = None
    if atlas_spec is not None:
        boundary_tides = _atlas_tides(case, atlas_spec, base, origin, start, notes)
    out = doc.get("output", {})
    _check_keys("output", out)
    stations = []
    for s in out.get("station", []):
        _check_keys("output.station", s)
        stations.append(Station(s["name"], float(s["x_m"]), float(s["y_m"]),
                                float(s.get("open_water_depth_m", DEFAULT_OPEN_WATER_DEPTH)),
                                float(s.get("max_offset_m", DEFAULT_MAX_OFFSET))))
    return CaseRun(run.get("name", "case"), case, 3600.0 * float(run.get("duration_hours", 1.0)),
                   60.0 * float(out.get("interval_minutes", 60.0)), stations, num.get("dt_max_s"),
                   int(num.get("local_time_stepping_levels", 0)), start, doc, boundary_tides, notes)


def _atlas_tides(case, spec: dict, base: str, origin, start, notes: list):
    """`[boundary.atlas]` as the reference's domain and runner build it (domain.rs:193-243,
    run.rs:706-730): read, corrected, its coverage checked (uncovered open faces become
    walls, within `max_closed`), projected onto the open boundary's nodes with the nodal
    terms of `run.start`, ramped and transport-scaled."""
    if origin is None:
        raise CaseError("[boundary.atlas] needs a placed mesh ([mesh] origin)")
    if start is None:
        raise CaseError("[boundary.atlas] needs run.start (the nodal corrections' epoch)")
    coverage = float(spec.get("coverage_m", 5000.0))
    max_closed = float(spec.get("max_closed", 0.05))
    ramp_hours = float(spec.get("ramp_hours", 1.0))
    ratio = spec.get("transport_scaling")
    if not coverage > 0.0:
        raise CaseError("boundary.atlas.coverage_m must be positive")
    if not 0.0 <= max_closed <= 1.0:
        raise CaseError("boundary.atlas.max_closed must be in [0, 1]")
    if ratio is not None and not float(ratio) > 0.0:
        raise CaseError("boundary.atlas.transport_scaling must be positive")
    if ramp_hours < 0.0:
        raise CaseError("boundary.atlas.ramp_hours must not be negative")
    path = spec["file"] if os.path.isabs(spec["file"]) else os.path.join(base, spec["file"])
    try:
        atlas = A.TidalAtlas.read(path)
        correction = spec.get("correction", "none")
        if correction == "minors":
            added = atlas.infer_missing_minors()
            notes.append(f"tidal atlas {path}: {len(added)} minor constituents inferred: {' '.join(added)}")
        elif correction == "theory":
            atlas.infer_from_theory()
        elif correction != "none":
            raise CaseError(f"[boundary.atlas] correction = {correction!r}: 'none', 'minors' or 'theory'")
        projection = A.LocalProjection(*origin)
        faces, length, missed = A.uncovered_faces(case, projection, atlas, coverage)
        closed = A.close_uncovered(case, faces, length, missed, max_closed)
        notes.append(f"tidal atlas {path}: {closed} open faces farther than {coverage / 1000.0:.1f} km made walls")
        clock = T.ModelClock.at_datetime(*start)
        return A.boundary_tides(case, atlas, projection, clock, coverage, ramp=3600.0 * ramp_hours,
                                max_ratio=float(ratio) if ratio is not None else 0.0)
    except (A.TidalAtlasError, OSError) as e:
        raise CaseError(f"{path}: {e}") from e


# ---------------------------------------------------------------------------
# Stations: a DG field at a point (Probe2D, src/solver/probe.rs; the point
# locator, src/mesh/core/point_locator.rs)
# ---------------------------------------------------------------------------
def inverse_bilinear(v, p):
    """(r, s) of p under the bilinear map of the quadrilateral v[4][2], Newton
    from the centre (point_locator.rs:60-96); None if it does not converge."""
    (x0, y0), (x1, y1), (x2, y2), (x3, y3) = v
    ax, ay = 0.25 * (x0 + x1 + x2 + x3), 0.25 * (y0 + y1 + y2 + y3)
    bx, by = 0.25 * (-x0 + x1 + x2 - x3), 0.25 * (-y0 + y1 + y2 - y3)
    cx, cy = 0.25 * (-x0 - x1 + x2 + x3), 0.25 * (-y0 - y1 + y2 + y3)
    dx, dy = 0.25 * (x0 - x1 + x2 - x3), 0.25 * (y0 - y1 + y2 - y3)
    noise = 16.0 * 2.220446049250313e-16 * max(abs(p[0]), abs(p[1]), abs(ax), abs(ay))
    r = s = 0.0
    for _ in range(50):
        fx = ax + bx * r + cx * s + dx * r * s - p[0]
        fy = ay + by * r + cy * s + dy * r * s - p[1]
        xr, xs = bx + dx * s, cx + dx * r
        yr, ys = by + dy * s, cy + dy * r
        det = xr * ys - xs * yr
        if det == 0.0 or not math.isfinite(det):
            return None
        dr = (ys * fx - xs * fy) / det
        ds = (-yr * fx + xr * fy) / det
        r -= dr
        s -= ds
        if max(abs(dr), abs(ds)) < 1e-14 * (1.0 + max(abs(r), abs(s))) + noise / math.sqrt(abs(det)):
            return r, s
        if max(abs(r), abs(s)) > 1e6:
            return None
    return None


class PointLocator:
    """The element holding a point: a bucket grid over the mesh's bounding box,
    about one bucket per element, each listing the elements whose bounding box
    overlaps it in element order; the first of a point's bucket that holds it
    (point_locator.rs:98-356)."""

    def __init__(self, mesh):
        self.corners = np.asarray(mesh.corners, dtype=np.float64)
        K = self.corners.shape[0]
        lo = self.corners.min(axis=1)
        hi = self.corners.max(axis=1)
        self.boxes = np.stack([lo[:, 0], hi[:, 0], lo[:, 1], hi[:, 1]], axis=1)
        x0, x1 = float(lo[:, 0].min()), float(hi[:, 0].max())
        y0, y1 = float(lo[:, 1].min()), float(hi[:, 1].max())
        width, height = max(x1 - x0, 0.0), max(y1 - y0, 0.0)
        n = float(max(K, 1))
        cell = math.sqrt(width * height / n)
        if not cell > 0.0:
            cell = max(width, height, 1.0) / n
        buckets = lambda extent: min(max(int(math.ceil(extent / cell)), 1), 4 * int(n))
        self.dims = (buckets(width), buckets(height))
        self.origin = (x0, y0)
        self.inv_cell = (self.dims[0] / max(width, cell), self.dims[1] / max(height, cell))
        self.lists = {}
        for k in range(K):
            b = self.boxes[k]
            i0, j0 = self.bucket(b[0], b[2])
            i1, j1 = self.bucket(b[1], b[3])
            for j in range(j0, j1 + 1):
                for i in range(i0, i1 + 1):
                    self.lists.setdefault(j * self.dims[0] + i, []).append(k)

    def bucket(self, x, y):
        def index(v, axis):
            f = math.floor((v - self.origin[axis]) * self.inv_cell[axis])
            return min(int(max(f, 0.0)), self.dims[axis] - 1)
        return index(x, 0), index(y, 1)

    def in_element(self, k, p):
        x0, x1, y0, y1 = self.boxes[k]
        slack = INSIDE_TOLERANCE * max(x1 - x0, y1 - y0)
        if p[0] < x0 - slack or p[0] > x1 + slack or p[1] < y0 - slack or p[1] > y1 + slack:
            return None
        rs = inverse_bilinear(self.corners[k], p)
        if rs is None:
            return None
        r, s = rs
        if abs(r) <= 1.0 + INSIDE_TOLERANCE and abs(s) <= 1.0 + INSIDE_TOLERANCE:
            return k, min(max(r, -1.0), 1.0), min(max(s, -1.0), 1.0)
        return None

    def locate(self, p):
        if not (math.isfinite(p[0]) and math.isfinite(p[1])):
            return None
        i, j = self.bucket(p[0], p[1])
        for k in self.lists.get(j * self.dims[0] + i, []):
            found = self.in_element(k, p)
            if found is not None:
                return found
        return None


def lagrange_basis(nodes, x):
    """l_i(x) = prod_{j != i} (x - x_j) / (x_i - x_j) (lagrange.rs:19-29)."""
    out = []
    for i, xi in enumerate(nodes):
        v = 1.0
        for j, xj in enumerate(nodes):
            if j != i:
                v *= (x - xj) / (xi - xj)
        out.append(v)
    return out


class Probe:
    """A point of one element and the nodal basis there (Probe2D)."""

    def __init__(self, element, r, s, order):
        self.element = element
        nodes = [float(v) for v in cases.gll_nodes(order)]
        lr, ls = lagrange_basis(nodes, r), lagrange_basis(nodes, s)
        # node j*n1 + i: l_i(r) l_j(s) (operators_2d.rs:253-266)
        self.weights = [a * b for b in ls for a in lr]

    def evaluate(self, values):
        acc = 0.0
        for w, v in zip(self.weights, values):
            acc += w * v
        return acc

    def sample(self, h, hu, hv, bed, nn):
        """Probe2D::sample_swe (probe.rs:112-140) with H_VELOCITY."""
        k = self.element
        sl = slice(k * nn, (k + 1) * nn)
        hh, qu, qv, b = (self.evaluate(a[sl]) for a in (h, hu, hv, bed))
        u, v = (qu / hh, qv / hh) if hh > H_VELOCITY else (0.0, 0.0)
        return {"h": hh, "hu": qu, "hv": qv, "bed": b, "eta": hh + b, "u": u, "v": v}


def station_probes(run: CaseRun):
    """Where each station samples (CaseDomain::stations, domain.rs:630-700): in
    its own element when that is open water (every bed node at least
    open_water_depth_m deep); a station whose element is not, and the nearest
    open element search that places it, are refused (not built)."""
    c = run.case
    loc = PointLocator(c.mesh)
    nn = c.nn
    out = []
    for s in run.stations:
        found = loc.locate((s.x, s.y))
        if found is None:
            raise CaseError(f"station {s.name!r} at ({s.x:.1f}, {s.y:.1f}) m is outside the mesh")
        k, r, sr = found
        if s.open_water_depth > 0.0 and not np.all(c.bed[k * nn:(k + 1) * nn] <= -s.open_water_depth):
            raise CaseError(f"station {s.name!r}: its element is shallower than {s.open_water_depth} m; the search "
                            "for the nearest open-water element is not built")
        out.append((s.name, Probe(k, r, sr, c.order)))
    return out


def rust_display(x: float) -> str:
    """A float as Rust's `{}` prints it: the shortest digits that round-trip,
    never in exponent form ("5", "0.1", "0.0000001", "-0", "NaN", "inf")."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    r = repr(abs(x))
    if "e" in r or "E" in r:
        mant, exp = r.lower().split("e")
        exp = int(exp)
        if "." in mant:
            ip, fp = mant.split(".")
        else:
            ip, fp = mant, ""
        digits = (ip + fp).lstrip("0") or "0"
        point = len(ip) + exp                 # position of the decimal point in `digits`
        if ip.strip("0") == "":
            lead = len(fp) - len(fp.lstrip("0"))
            digits = fp.lstrip("0")
            point = exp - lead
        if point <= 0:
            s = "0." + "0" * (-point) + digits
        elif point >= len(digits):
            s = digits + "0" * (point - len(digits))
        else:
            s = digits[:point] + "." + digits[point:]
    else:
        s = r[:-2] if r.endswith(".0") else r
    return sign + s


# ---------------------------------------------------------------------------
# The runner's restart data (src/case/run/restart.rs:1-20,191-275)
# ---------------------------------------------------------------------------
BUILD = "dgcom_amd (MI355X HIP path)"
STATION_FIELDS = ("time_s", "eta_m", "u_m_s", "v_m_s", "depth_m", "hs_m", "tp_s", "wave_from_deg")


def _toml_value(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, float)):
        return repr(v) if not (isinstance(v, float) and v == int(v) and abs(v) < 1e16) else f"{v:.1f}"
    if isinstance(v, str):
        return '"' + v.replace("\\", "\\\\").replace('"', '\\"') + '"'
    if isinstance(v, (_dt.datetime, _dt.date)):
        return v.isoformat()
    if isinstance(v, list):
        return "[" + ", ".join(_toml_value(x) for x in v) + "]"
    raise CaseError(f"a case value of type {type(v).__name__} cannot be written")


def to_toml(doc: dict) -> str:
    """The case as TOML text: a table's plain keys, then its sub-tables and
    arrays of tables (`[[boundary]]`), depth first."""
    out = []

    def walk(prefix, table):
        plain = {k: v for k, v in table.items()
                 if not isinstance(v, dict) and not (isinstance(v, list) and v and all(isinstance(x, dict) for x in v))}
        for k, v in plain.items():
            out.append(f"{k} = {_toml_value(v)}")
        for k, v in table.items():
            name = f"{prefix}.{k}" if prefix else k
            if isinstance(v, dict):
                out.append(f"\n[{name}]")
                walk(name, v)
            elif isinstance(v, list) and v and all(isinstance(x, dict) for x in v):
                for item in v:
                    out.append(f"\n[[{name}]]")
                    walk(name, item)

    walk("", doc)
    return "\n".join(out).lstrip("\n") + "\n"


def driver_data(run: CaseRun, records, volume_start: float, steps: int):
    """A case run's restart data in the reference runner's layout
    (restart.rs:191-213): metadata `build` and each station's name
    `station.<j>`; the text `case` (the case with its build line); the arrays
    `output.counters` (VTK frames written: none here; the volume at the start)
    and `station.<j>.<field>` (its records so far; the sea-state fields empty
    without waves). `steps` is this runner's own addition (metadata the
    reference does not read)."""
    metadata = [("build", BUILD)] + [(f"station.{j}", s.name) for j, s in enumerate(run.stations)]
    metadata.append(("steps", str(steps)))
    text = [("case", f"# {BUILD}\n{to_toml(run.doc)}")]
    extra = [("output.counters", np.array([0.0, volume_start]))]
    names = [s.name for s in run.stations]
    for j, name in enumerate(names):
        rows = np.array([r[1:] for r in records if r[0] == name], dtype=np.float64).reshape(-1, 5)
        for c, field in enumerate(STATION_FIELDS[:5]):
            extra.append((f"station.{j}.{field}", rows[:, c].copy()))
        for field in STATION_FIELDS[5:]:
            extra.append((f"station.{j}.{field}", np.zeros(0)))
    return metadata, text, extra


def restore_driver_data(run: CaseRun, rst):
    """The records, the volume at the start and the steps so far from a
    restart's driver data (restart.rs:216-275): refused by name without
    `output.counters`, with other stations, a missing record array or more
    stations than the run; the records interleaved by time as recorded."""
    extra = dict(rst.extra)
    meta = dict(rst.metadata)
    counters = extra.get("output.counters")
    if counters is None or counters.size != 2:
        raise CaseError("the restart has no output.counters")
    series = []
    for j, s in enumerate(run.stations):
        theirs = meta.get(f"station.{j}")
        if theirs != s.name:
            raise CaseError(f"the restart's stations: station {j} is {s.name}, the restart's {theirs or '(none)'}")
        cols = []
        for field in STATION_FIELDS:
            v = extra.get(f"station.{j}.{field}")
            if v is None:
                raise CaseError(f"the restart's stations: no station.{j}.{field}")
            cols.append(v)
        if any(c.size != cols[0].size for c in cols[1:5]):
            raise CaseError(f"the restart's stations: station {j}'s records differ in length")
        series.append(cols)
    if f"station.{len(run.stations)}" in meta:
        raise CaseError("the restart's stations: it has more stations than this run")
    records = []
    n = max((c[0].size for c in series), default=0)
    for i in range(n):
        for s, cols in zip(run.stations, series):
            if i < cols[0].size:
                records.append((s.name, *(float(c[i]) for c in cols[:5])))
    return records, float(counters[1]), int(meta.get("steps", "0"))


# ---------------------------------------------------------------------------
# The runner (case::run, src/case/run.rs; Simulation::run_with_callback)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class RunReport:
    t: float
    steps: int
    records: list       # [(station, t, eta, u, v, depth)]
    clips: int
    state: tuple = ()   # (h, hu, hv) at the end
    volume: tuple = ()  # the volume at the run's start (time 0) and at its end, m^3 (run.rs:626)


def run_case(run: CaseRun, out_dir: str | None = None, device: int = 0, restart_from: str | None = None,
             stop_at: float | None = None, restart_to: str | None = None, physics=None) -> RunReport:
    """Run a case on the GPU from still water (or from a restart) to its end
    (or `stop_at`), landing on every output time, recording the stations there
    (and at the start of a fresh run) into `out_dir/stations.csv`, and writing a
    restart at the end if asked (its driver data carry the records so far and
    the counters, in the reference runner's layout)."""
    from . import swe2d
    c = run.case
    phys = physics or swe2d.SWEPhysics2D.from_case(c, device=device)
    if run.boundary_tides is not None:
        phys.set_boundary_tides(run.boundary_tides)
    masses = cases.node_masses(c)
    t = 0.0
    h0, hu0, hv0 = c.h, c.hu, c.hv
    records = []
    steps = 0
    if restart_from is not None:
        rst = restart.Restart2D.read(restart_from)
        rst.check_domain(c.mesh, c.bed, c.order)
        t, (h0, hu0, hv0) = rst.time, (rst.h, rst.hu, rst.hv)
        records, volume_start, steps = restore_driver_data(run, rst)
    else:
        volume_start = float(np.sum(masses * h0))
    state = phys.state(h0, hu0, hv0)
    probes = station_probes(run)
    t_end = run.duration_s if stop_at is None else min(stop_at, run.duration_s)

    def record(x, time):
        h, hu, hv = x
        for name, pr in probes:
            smp = pr.sample(h, hu, hv, c.bed, c.nn)
            records.append((name, time, smp["eta"], smp["u"], smp["v"], smp["h"]))

    if restart_from is None:
        record(state.download(), t)
    mr = swe2d.Multirate(phys, run.lts_levels, c.scheme) if run.lts_levels > 0 else None
    k = int(math.floor(t / run.interval_s + 0.5))
    while t < t_end:
        target = min((k + 1) * run.interval_s, t_end)
        if mr is not None:
            fitted, dt = mr.assign_levels(state, c.cfl, run.dt_max, target - t)
            if run.dt_max is not None:
                dt = min(dt, run.dt_max)
            dt = min(dt, fitted)
        else:
            dt = phys.compute_dt_ssp(state, c.cfl, c.scheme)
            if run.dt_max is not None:
                dt = min(dt, run.dt_max)
        landed = t + dt * (1.0 + LANDING_SLACK) >= target
        if landed:
            dt = target - t
        if mr is not None:
            mr.step(state, t, dt)
        else:
            phys.step(state, t, dt, c.scheme)
        t = target if landed else t + dt
        steps += 1
        if not state.is_finite():
            raise CaseError(f"The state is not finite at t = {t:.1f} s: it has blown up")
        if landed:
            # the runner records at every output time and at the run's end (run.rs:1420-1440)
            on_output = target == (k + 1) * run.interval_s
            if on_output:
                k += 1
            if on_output or t == run.duration_s:
                record(state.download(), t)
    x = state.download()
    if mr is not None:
        mr.close()
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)
        write_stations(os.path.join(out_dir, "stations.csv"), records)
    if restart_to is not None:
        metadata, text, extra = driver_data(run, records, volume_start, steps)
        restart.Restart2D(t, restart.domain_fingerprint(c.mesh, c.bed), x[0], x[1], x[2], c.n_elements, c.nn,
                          metadata=metadata, text=text, extra=extra).write(restart_to)
    clips = phys.negative_depth_clips()
    state.free()
    return RunReport(t, steps, records, clips, x, (volume_start, float(np.sum(masses * x[0]))))


def write_stations(path: str, records):
    """stations.csv as the runner writes it (run.rs:1320-1335,1410-1418)."""
    with open(path, "w") as f:
        f.write("station,time_s,eta_m,u_m_s,v_m_s,depth_m\n")
        for name, t, eta, u, v, depth in records:
            f.write(",".join([name] + [rust_display(float(x)) for x in (t, eta, u, v, depth)]) + "\n")
