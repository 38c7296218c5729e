"""Workload generators: quadrilateral meshes, beds, forcing and initial states.

Everything here is plain input data (numpy), shared by the HIP path and by its
checker. The mesh connectivity uses the reference's 4-byte face link
(`src/mesh/core/mesh2d.rs:54-107`): ``element << 4 | shared << 3 | left << 2 |
face`` for an interior face, ``0xFFFFFFFF`` for a boundary face, whose
boundary kind sits in a separate byte per face (wall 0, tidal open boundary 1).

Structured grids follow `Mesh2D::uniform_rectangle_with_sides` /
`uniform_periodic` (`mesh2d.rs:216-243,309-400`): vertices row by row,
element ``k = j*nx + i`` counter-clockwise from its south-west corner, faces
0..3 = south, east, north, west; the left side of an edge is the face that
appears first in (element, face) order, and of a periodic pair the east /
north face.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BOUNDARY = 0xFFFFFFFF
WALL = 0
TIDE = 1

# M2 period of the workload definitions: 12.42 h. (The reference's own
# Doodson-number M2 is 12.4206012 h, `src/tides/astronomy.rs:360-363`.)
M2_PERIOD_S = 12.42 * 3600.0


def gll_nodes(order: int) -> np.ndarray:
    """Closed-form Gauss-Lobatto-Legendre points on [-1, 1] for N = 1..3."""
    if order == 1:
        return np.array([-1.0, 1.0])
    if order == 2:
        return np.array([-1.0, 0.0, 1.0])
    if order == 3:
        a = math.sqrt(0.2)
        return np.array([-1.0, -a, a, 1.0])
    raise ValueError("order must be 1, 2 or 3")


def link(element: int, face: int, left: bool, shared: bool) -> int:
    return (element << 4) | (int(shared) << 3) | (int(left) << 2) | face


@dataclasses.dataclass
class Mesh:
    corners: np.ndarray  # (K, 4, 2) float64, counter-clockwise from (r, s) = (-1, -1)
    links: np.ndarray    # (K, 4) uint32
    tags: np.ndarray     # (K, 4) uint8, boundary kind of boundary faces
    vertices: np.ndarray = None  # (V, 2) the shared vertices, where the builder keeps them
    quads: np.ndarray = None     # (K, 4) each element's vertex indices

    @property
    def n_elements(self) -> int:
        return self.corners.shape[0]


def structured_mesh(x0, x1, y0, y1, nx, ny, sides=(WALL, WALL, WALL, WALL), periodic_x=False,
                    periodic_y=False, jitter=0.0, seed=0, distort=None) -> Mesh:
    """An nx x ny grid of quads. ``sides`` = boundary kinds of [south, east, north,
    west]. ``jitter``: interior vertices moved by independent uniform offsets of up
    to ``jitter`` times the cell size in x and y (seeded). ``distort``: a callable
    (x, y) -> (x', y') applied to every vertex (smooth periodic distortions)."""
    dx = (x1 - x0) / nx
    dy = (y1 - y0) / ny
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    vx = x0 + ii * dx
    vy = y0 + jj * dy
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        ox = rng.uniform(-jitter, jitter, size=vx.shape) * dx
        oy = rng.uniform(-jitter, jitter, size=vy.shape) * dy
        interior = np.zeros(vx.shape, dtype=bool)
        interior[1:-1, 1:-1] = True
        vx = np.where(interior, vx + ox, vx)
        vy = np.where(interior, vy + oy, vy)
    if distort is not None:
        vx, vy = distort(vx, vy)
    K = nx * ny
    j, i = np.divmod(np.arange(K), nx)
    corners = np.empty((K, 4, 2))
    for c, (di, dj) in enumerate(((0, 0), (1, 0), (1, 1), (0, 1))):
        corners[:, c, 0] = vx[j + dj, i + di]
        corners[:, c, 1] = vy[j + dj, i + di]
    links = np.full((K, 4), BOUNDARY, dtype=np.uint64)
    tags = np.zeros((K, 4), dtype=np.uint8)
    k = np.arange(K, dtype=np.uint64)

    def put(mask, face, nb, nb_face, left, shared):
        links[mask, face] = (nb[mask] << 4) | (int(shared) << 3) | (int(left) << 2) | nb_face

    # south (face 0) <-> north (face 2) of the element below; east (1) <-> west (3)
    put(j > 0, 0, k - nx, 2, False, True)
    put(j < ny - 1, 2, k + nx, 0, True, True)
    put(i < nx - 1, 1, k + 1, 3, True, True)
    put(i > 0, 3, k - 1, 1, False, True)
    if periodic_x:
        put(i == nx - 1, 1, k - (nx - 1), 3, True, False)
        put(i == 0, 3, k + (nx - 1), 1, False, False)
    if periodic_y:
        put(j == ny - 1, 2, k - (ny - 1) * nx, 0, True, False)
        put(j == 0, 0, k + (ny - 1) * nx, 2, False, False)
    for face, (mask) in enumerate((j == 0, i == nx - 1, j == ny - 1, i == 0)):
        b = mask & (links[:, face] == BOUNDARY)
        tags[b, face] = sides[face]
    # the grid's vertices, row by row, as the reference's structured_quads numbers them (mesh2d.rs:216-243)
    v0 = j * (nx + 1) + i
    quads = np.stack([v0, v0 + 1, v0 + nx + 2, v0 + nx + 1], axis=1).astype(np.uint64)
    vertices = np.stack([vx.ravel(), vy.ravel()], axis=1)
    return Mesh(corners=corners, links=links.astype(np.uint32), tags=tags, vertices=vertices, quads=quads)


def node_coordinates(mesh: Mesh, order: int):
    """Physical (x, y) of every GLL node, (K, NN) each, node i = j*(N+1) + a
    (r fastest), by the bilinear map of `Mesh2D::bilinear` (`mesh2d.rs:578-595`)."""
    x1 = gll_nodes(order)
    s, r = np.meshgrid(x1, x1, indexing="ij")
    r = r.ravel()
    s = s.ravel()
    n0 = (1.0 - r) * (1.0 - s) / 4.0
    n1 = (1.0 + r) * (1.0 - s) / 4.0
    n2 = (1.0 + r) * (1.0 + s) / 4.0
    n3 = (1.0 - r) * (1.0 + s) / 4.0
    c = mesh.corners
    x = n0 * c[:, 0, 0:1] + n1 * c[:, 1, 0:1] + n2 * c[:, 2, 0:1] + n3 * c[:, 3, 0:1]
    y = n0 * c[:, 0, 1:2] + n1 * c[:, 1, 1:2] + n2 * c[:, 2, 1:2] + n3 * c[:, 3, 1:2]
    return x, y


def ramp_depth(x, y, from_m, to_m, depth_m):
    """`Ramp::depth` (`src/case/schema.rs:656-663`)."""
    x0, y0 = from_m
    dx, dy = to_m[0] - x0, to_m[1] - y0
    s = np.clip(((x - x0) * dx + (y - y0) * dy) / (dx * dx + dy * dy), 0.0, 1.0)
    return depth_m[0] + s * (depth_m[1] - depth_m[0])


def gaussian_random_field(x, y, extent, sd, correlation, seed, spacing=None):
    """A seeded, zero-mean Gaussian random field with a squared-exponential
    covariance ``sd^2 exp(-r^2 / (2 correlation^2))``, by spectral synthesis on a
    periodic grid twice the domain, sampled at (x, y) by bilinear interpolation.
    ``extent`` = (x0, x1, y0, y1)."""
    x0, x1, y0, y1 = extent
    size = max(x1 - x0, y1 - y0) * 2.0
    if spacing is None:
        spacing = correlation / 8.0
    m = int(2 ** math.ceil(math.log2(size / spacing)))
    h = size / m
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((m, m))
    k = np.fft.fftfreq(m, d=h) * 2.0 * math.pi
    kx, ky = np.meshgrid(k, k, indexing="ij")
    psd = np.exp(-0.5 * (kx * kx + ky * ky) * correlation * correlation)
    field = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.sqrt(psd)))
    field *= sd / field.std()
    gx = (np.asarray(x) - x0) / h
    gy = (np.asarray(y) - y0) / h
    ix = np.floor(gx).astype(np.int64)
    iy = np.floor(gy).astype(np.int64)
    fx = gx - ix
    fy = gy - iy
    ix %= m
    iy %= m
    ix1 = (ix + 1) % m
    iy1 = (iy + 1) % m
    return ((1 - fx) * (1 - fy) * field[ix, iy] + fx * (1 - fy) * field[ix1, iy]
            + (1 - fx) * fy * field[ix, iy1] + fx * fy * field[ix1, iy1])


@dataclasses.dataclass
class Tide:
    """Uniform harmonic tide `eta(t) = mean + R(t) sum f A cos(w t + phase + (V0 + u)
    pi/180)`, with the smooth ramp `R = 3 tau^2 - 2 tau^3`, `tau = t / ramp`
    (`harmonic_tide.rs:24-32,124-136`). `nodal`: per constituent (f0, f_rate, V0,
    u0, u_rate) of its nodal drift (degrees, seconds; `tides.ModelClock.nodal_drift`),
    empty for none (f = 1, V0 = u = 0)."""
    amplitude: tuple = ()
    omega: tuple = ()
    phase: tuple = ()
    mean: float = 0.0
    ramp: float = 3600.0
    nodal: tuple = ()

    def elevation(self, t: float) -> float:
        tide = 0.0
        for i, (a, w, p) in enumerate(zip(self.amplitude, self.omega, self.phase)):
            f0, f_rate, v0, u0, u_rate = self.nodal[i] if self.nodal else (1.0, 0.0, 0.0, 0.0, 0.0)
            f = f0 + f_rate * t
            u = u0 + u_rate * t
            tide += f * a * math.cos(w * t + p + (v0 + u) * (math.pi / 180.0))
        rf = 1.0
        if self.ramp > 0.0:
            tau = min(max(t / self.ramp, 0.0), 1.0)
            rf = tau * tau * (3.0 - 2.0 * tau)
        return self.mean + rf * tide


def m2_tide(amplitude, phase=0.0, ramp=3600.0) -> Tide:
    return Tide(amplitude=(amplitude,), omega=(2.0 * math.pi / M2_PERIOD_S,), phase=(phase,), ramp=ramp)


@dataclasses.dataclass
class Case:
    """One 2D shallow-water problem: mesh, order, bed, friction, rotation, tide,
    initial state and stepping."""
    name: str
    order: int
    mesh: Mesh
    bed: np.ndarray            # (K*NN,)
    manning: np.ndarray        # (K,) per-element Manning n
    f0: float
    tide: Tide
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    cfl: float = 0.5
    scheme: int = 0            # 0 SSP-RK3, 1 SSP-RK(4,3)
    n_steps: int = 1000
    g: float = 9.81
    h_min: float = 1e-6
    h_dry: float = 1e-3
    incoming_wave: bool = False  # tide data as a wave entering from still water, not at rest

    @property
    def n_elements(self) -> int:
        return self.mesh.n_elements

    @property
    def nn(self) -> int:
        return (self.order + 1) ** 2

    def dofs(self) -> int:
        return 3 * self.n_elements * self.nn


def still_water(bed, eta0=0.0):
    """`still_water` (`src/case/run.rs:1158-1168`): h = max(0, eta0 - B), at rest."""
    h = np.maximum(eta0 - bed, 0.0)
    return h, np.zeros_like(h), np.zeros_like(h)


def tidal_estuary(nx=100, ny=20, order=2, n_steps=1000) -> Case:
    """Idealised tidal estuary: 10 km x 2 km, 100 x 20 affine quads, N = 2, bed depth
    linear from 10 m at the open (west) end to 2 m at the closed end, 0.5 m M2 tide
    on the west boundary, Manning 0.025, f = 1.2e-4, SSP-RK3 at CFL 0.5, from rest."""
    mesh = structured_mesh(0.0, 10000.0, 0.0, 2000.0, nx, ny, sides=(WALL, WALL, WALL, TIDE))
    x, y = node_coordinates(mesh, order)
    bed = -ramp_depth(x, y, (0.0, 1000.0), (10000.0, 1000.0), (10.0, 2.0)).ravel()
    h, hu, hv = still_water(bed)
    return Case("tidal_estuary", order, mesh, bed, np.full(mesh.n_elements, 0.025), 1.2e-4, m2_tide(0.5),
                h, hu, hv, cfl=0.5, scheme=0, n_steps=n_steps)


def large_unstructured(n=1000, length=100_000.0, order=3, seed=1234, n_steps=1000, correlation=5000.0,
                       mean_depth=40.0, sd=25.0, amplitude=1.0) -> Case:
    """Large unstructured: an n x n grid over `length` square with interior vertices
    jittered by up to 20 % of the cell size (seeded), N = 3, bed = -(mean_depth +
    GRF(sd, correlation)) so part of the domain is dry or intertidal, an M2 tide on
    the west boundary, Manning 0.025, f = 1.3e-4, SSP-RK3. The defaults are the
    1,000,000-element workload; smaller `n` keeps the same physics."""
    mesh = structured_mesh(0.0, length, 0.0, length, n, n, sides=(WALL, WALL, WALL, TIDE), jitter=0.2, seed=seed)
    x, y = node_coordinates(mesh, order)
    field = gaussian_random_field(x, y, (0.0, length, 0.0, length), sd, correlation, seed + 1)
    bed = -(mean_depth + field).ravel()
    h, hu, hv = still_water(bed)
    return Case("large_unstructured", order, mesh, bed, np.full(mesh.n_elements, 0.025), 1.3e-4, m2_tide(amplitude),
                h, hu, hv, cfl=0.5, scheme=0, n_steps=n_steps)


def drying_beach(nx=24, ny=6, order=2, slope=0.01, length=2000.0, width=500.0, eta_amp=0.5, jitter=0.0,
                 seed=7, n_steps=200) -> Case:
    """A walled sloping beach (deep at the open west end, rising above the mean
    level to the east) forced by a tide: it wets and dries, exercising the
    subcells, the wet/dry correction and the tidal boundary."""
    mesh = structured_mesh(0.0, length, 0.0, width, nx, ny, sides=(WALL, WALL, WALL, TIDE), jitter=jitter, seed=seed)
    x, y = node_coordinates(mesh, order)
    bed = (-5.0 + slope * x + 0.3 * np.sin(2 * math.pi * y / width)).ravel()
    h, hu, hv = still_water(bed)
    tide = Tide(amplitude=(eta_amp,), omega=(2 * math.pi / 600.0,), phase=(0.0,), ramp=60.0)
    return Case("drying_beach", order, mesh, bed, np.full(mesh.n_elements, 0.025), 1.0e-4, tide, h, hu, hv,
                cfl=0.5, scheme=0, n_steps=n_steps)



def ensemble_members(n_members=64, seed=99, k_lo=20_000, k_hi=400_000, n_steps=500) -> list:
    """The batched forward ensemble: seeded members with K log-uniform in [k_lo, k_hi]
    jittered quads of 100 m, N uniform in {1, 2, 3}, an M2 tide of amplitude uniform
    in 0.2-1.5 m and phase in 0-2pi, and their own Gaussian-random-field beds."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n_members):
        K = float(np.exp(rng.uniform(np.log(k_lo), np.log(k_hi))))
        n = max(4, int(round(np.sqrt(K))))
        order = int(rng.integers(1, 4))
        amp, phase = float(rng.uniform(0.2, 1.5)), float(rng.uniform(0.0, 2.0 * np.pi))
        c = large_unstructured(n=n, length=100.0 * n, order=order, seed=seed + 1000 + 17 * m, n_steps=n_steps)
        c.tide = Tide(amplitude=(amp,), omega=(2.0 * np.pi / M2_PERIOD_S,), phase=(phase,), ramp=3600.0)
        out.append(c)
    return out

# ---------------------------------------------------------------------------
# Thacker's planar surface in a paraboloid (tests/wet_dry_2d_test.rs:217-310)
# ---------------------------------------------------------------------------
THACKER_A, THACKER_H0, THACKER_ETA, THACKER_L = 1.0, 0.1, 0.5, 4.0   # SWASHES parameters


def thacker_omega(g=9.81):
    return math.sqrt(2.0 * g * THACKER_H0) / THACKER_A


def thacker_bed(x, y):
    r2 = (x - 0.5 * THACKER_L) ** 2 + (y - 0.5 * THACKER_L) ** 2
    return -THACKER_H0 * (1.0 - r2 / (THACKER_A * THACKER_A))


def thacker_exact(x, y, t, g=9.81):
    """The exact (h, hu, hv) at points (x, y) and time t: a planar surface whose
    wet disc slides round the bowl with period 2 pi / omega."""
    w = thacker_omega(g)
    xc, yc = x - 0.5 * THACKER_L, y - 0.5 * THACKER_L
    surface = THACKER_ETA * THACKER_H0 / (THACKER_A * THACKER_A) * (
        2.0 * xc * math.cos(w * t) + 2.0 * yc * math.sin(w * t) - THACKER_ETA)
    h = surface - thacker_bed(x, y)
    wet = h > 0.0
    u, v = -THACKER_ETA * w * math.sin(w * t), THACKER_ETA * w * math.cos(w * t)
    h = np.where(wet, h, 0.0)
    return h, np.where(wet, h * u, 0.0), np.where(wet, h * v, 0.0)


def thacker(n=40, order=2, h_dry=1e-5) -> Case:
    """Thacker's case on an n x n mesh of [0, 4] m^2 (walls, no friction, no
    rotation), started from the exact state at t = 0, SSP-RK3 at CFL 1 for one
    period; dry threshold 1e-5 m (1e-4 of the 0.1 m bowl depth, as the
    reference's test takes it)."""
    mesh = structured_mesh(0.0, THACKER_L, 0.0, THACKER_L, n, n)
    x, y = node_coordinates(mesh, order)
    bed = thacker_bed(x, y).ravel()
    h, hu, hv = thacker_exact(x.ravel(), y.ravel(), 0.0)
    return Case("thacker", order, mesh, bed, np.zeros(mesh.n_elements), 0.0, None, h, hu, hv, cfl=1.0, scheme=0,
                n_steps=0, h_dry=h_dry)


def node_masses(case) -> np.ndarray:
    """GLL quadrature weight times J at every node (K*NN,): the integral of a
    nodal field f is sum(node_masses * f)."""
    from numpy.polynomial import legendre as _leg
    n = case.order
    x1 = gll_nodes(n)
    pn = _leg.legval(x1, [0] * n + [1])
    w1 = 2.0 / (n * (n + 1) * pn * pn)
    w2 = np.outer(w1, w1).ravel()   # node i = j (N+1) + a: w_a w_j
    c = case.mesh.corners
    out = np.empty((case.n_elements, (n + 1) ** 2))
    for j in range(n + 1):
        for a in range(n + 1):
            r, s = x1[a], x1[j]
            xr = 0.25 * ((1 - s) * (c[:, 1, 0] - c[:, 0, 0]) + (1 + s) * (c[:, 2, 0] - c[:, 3, 0]))
            yr = 0.25 * ((1 - s) * (c[:, 1, 1] - c[:, 0, 1]) + (1 + s) * (c[:, 2, 1] - c[:, 3, 1]))
            xs = 0.25 * ((1 - r) * (c[:, 3, 0] - c[:, 0, 0]) + (1 + r) * (c[:, 2, 0] - c[:, 1, 0]))
            ys = 0.25 * ((1 - r) * (c[:, 3, 1] - c[:, 0, 1]) + (1 + r) * (c[:, 2, 1] - c[:, 1, 1]))
            out[:, j * (n + 1) + a] = w2[j * (n + 1) + a] * (xr * ys - xs * yr)
    return out.ravel()


# ---------------------------------------------------------------------------
# The reference's example case cases/tidal_channel/case.toml
# ---------------------------------------------------------------------------
EARTH_ROTATION_RATE = 7.292_115e-5   # source/swe_2d/coriolis.rs:25


def coriolis_parameter(latitude_deg):
    """f = 2 Omega sin(latitude) (coriolis.rs:29-31)."""
    return 2.0 * EARTH_ROTATION_RATE * math.sin(math.radians(latitude_deg))


def ramp_depth(x, y, from_m, to_m, depth_m):
    """Depth along a ramp, constant across it (case/schema.rs:656-663)."""
    x0, y0 = from_m
    dx, dy = to_m[0] - x0, to_m[1] - y0
    s = np.clip(((x - x0) * dx + (y - y0) * dy) / (dx * dx + dy * dy), 0.0, 1.0)
    return depth_m[0] + s * (depth_m[1] - depth_m[0])


def harmonic_tide(constituents, ramp_hours, start=None, mean=0.0):
    """The case runner's tide (case/run.rs:686-704): each (name, amplitude, Greenwich
    lag G in degrees) as A cos(w t - G) with w = 2 pi / the constituent's period
    from its Doodson speed, ramped over `ramp_hours`, with the nodal corrections of
    a run starting at `start` = (year, month, day[, hour, minute]) UTC if given."""
    from . import tides
    amp, omega, phase, nodal = [], [], [], []
    clock = tides.ModelClock.at_datetime(*start) if start is not None else None
    for name, a, g in constituents:
        amp.append(a)
        omega.append(2.0 * math.pi / tides.constituent_period(name))
        phase.append(-math.radians(g))
        if clock is not None:
            d = clock.nodal_drift(name)
            nodal.append((d.at_epoch.f, d.f_rate, d.at_epoch.v0_deg, d.at_epoch.u_deg, d.u_rate_deg))
    return Tide(amplitude=tuple(amp), omega=tuple(omega), phase=tuple(phase), mean=mean, ramp=3600.0 * ramp_hours,
                nodal=tuple(nodal))


def tidal_channel(order=2, nx=40, ny=4, velocity="at_rest") -> Case:
    """cases/tidal_channel/case.toml: a channel 40 km x 4 km open to the sea at its
    west end, its bed rising from 30 m below mean sea level at the mouth to 2 m
    above at the head (tidal flats); M2 1.0 m / G 0 and S2 0.35 m / G 30 deg with
    a 2 h ramp and the nodal corrections of a run from 2025-06-15 00:00 UTC through
    the characteristic boundary, its external velocity `velocity` ("at_rest", the
    case's, or "incoming_wave"); Coriolis at 63.5 N, Manning 0.025, wet/dry
    with the defaults; still water at mean sea level; SSP-RK(4,3) at its default
    CFL, 0.9 of the linear limit (1.359 at N = 2)."""
    mesh = structured_mesh(0.0, 40_000.0, 0.0, 4_000.0, nx, ny, sides=(WALL, WALL, WALL, TIDE))
    x, y = node_coordinates(mesh, order)
    bed = -ramp_depth(x, y, (0.0, 2000.0), (40_000.0, 2000.0), (30.0, -2.0)).ravel()
    h, hu, hv = still_water(bed)
    tide = harmonic_tide((("M2", 1.0, 0.0), ("S2", 0.35, 30.0)), 2.0, start=(2025, 6, 15))
    lin43 = (2.2, 1.51, 1.15, 0.93)   # linear CFL limits of SSP-RK(4,3) (swe_2d.rs:1263-1302)
    if velocity not in ("at_rest", "incoming_wave"):
        raise ValueError("velocity must be 'at_rest' or 'incoming_wave'")
    return Case("tidal_channel", order, mesh, bed, np.full(mesh.n_elements, 0.025), coriolis_parameter(63.5), tide,
                h, hu, hv, cfl=0.9 * lin43[order - 1], scheme=1, n_steps=0, incoming_wave=velocity == "incoming_wave")


# ---------------------------------------------------------------------------
# Local time stepping: the reference's graded channel (tests/local_time_stepping_test.rs:140-147,404-410)
# ---------------------------------------------------------------------------
def graded_mesh(lx, ly, nx, ny, ratio, sides=(WALL, WALL, WALL, WALL)) -> Mesh:
    """nx columns growing geometrically by `ratio` over [0, lx] x [0, ly]
    (graded_mesh, local_time_stepping_test.rs:140-147)."""
    n = float(nx)

    def grade(vx, vy):
        return lx * (ratio ** (n * vx) - 1.0) / (ratio ** n - 1.0), vy

    return structured_mesh(0.0, 1.0, 0.0, ly, nx, ny, sides=sides, distort=grade)


def _gaussian(x, y, x0, y0, width):
    return np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (width * width))


def graded_channel(order=2, nx=16, ny=3, ratio=1.25, hump=0.5, island=False, friction=False, scheme=0,
                   cfl=0.5) -> Case:
    """The reference's local-time-stepping channel (local_time_stepping_test.rs:
    404-499): 10 km x 2 km, `nx` columns growing by `ratio` (16 by 1.25: 28x from
    the first to the last), a bed from 20 to 110 m deep with a cross-channel
    cosine, walls all round, no rotation. Without `island`: a Gaussian hump of
    `hump` m at (2 km, 1 km) on still water (several_levels_conserve_mass). With
    `island`: a rough bed with an island above the surface, the lake at rest,
    Manning 0.025 (lake_at_rest_with_dry_land_across_levels)."""
    mesh = graded_mesh(10_000.0, 2_000.0, nx, ny, ratio)
    x, y = node_coordinates(mesh, order)
    x, y = x.ravel(), y.ravel()
    bed = -(20.0 + 90.0 * x / 10_000.0) - 5.0 * np.cos(2.0 * math.pi * y / 2_000.0)
    if island:
        bed = bed + 8.0 * np.sin(x / 700.0) * np.cos(y / 400.0) + 140.0 * _gaussian(x, y, 6_000.0, 1_000.0, 900.0)
        eta = np.zeros_like(x)
    else:
        eta = hump * _gaussian(x, y, 2_000.0, 1_000.0, 800.0)
    h = np.maximum(eta - bed, 0.0)
    return Case("graded_channel", order, mesh, bed, np.full(mesh.n_elements, 0.025 if friction else 0.0), 0.0,
                Tide(), h, np.zeros_like(h), np.zeros_like(h), cfl=cfl, scheme=scheme, n_steps=0)


# ---------------------------------------------------------------------------
# Tides per boundary node (BoundaryTides, src/boundary/tidal_atlas.rs:858-1140)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class BoundaryTides:
    """The open boundary's tides per boundary face node: constituents' angular
    speeds `omega` (rad/s, the nodal phase drift included) and nodal factor
    drifts `f_rate` (1/s); `slot` [K][4][N+1] (0xFFFFFFFF: the uniform tide);
    per slot `coef` [n_c][6] (C, S of eta, u, v: f(t)/f0 (C cos w't + S sin w't)),
    `mean` (added unramped) and `source_depth`; the ramp (s), whether the tides
    carry a velocity and the transport ratio limit (0: unscaled)."""
    omega: np.ndarray
    f_rate: np.ndarray
    slot: np.ndarray
    coef: np.ndarray
    mean: np.ndarray
    source_depth: np.ndarray
    ramp: float = 0.0
    velocity: bool = False
    max_ratio: float = 0.0


def face_node(n1, f, fi):
    """The element node of face f's node fi (faces counter-clockwise from (-1, -1))."""
    return (fi if f == 0 else fi * n1 + n1 - 1 if f == 1 else (n1 - 1) * n1 + (n1 - 1 - fi) if f == 2
            else (n1 - 1 - fi) * n1)


def open_face_nodes(case):
    """The open-boundary face nodes (k, f, fi) of a case, in mesh order, with
    their coordinates."""
    n1 = case.order + 1
    x, y = node_coordinates(case.mesh, case.order)
    x, y = x.reshape(case.n_elements, -1), y.reshape(case.n_elements, -1)
    fnodes = [[face_node(n1, f, fi) for fi in range(n1)] for f in range(4)]
    out = []
    for k in range(case.n_elements):
        for f in range(4):
            if case.mesh.links[k, f] == BOUNDARY and case.mesh.tags[k, f] == TIDE:
                for fi in range(n1):
                    nd = fnodes[f][fi]
                    out.append((k, f, fi, float(x[k, nd]), float(y[k, nd])))
    return out


def boundary_tides(case, omega, f_rate, constants, ramp=0.0, velocity=False, max_ratio=0.0, mean=None,
                   source_depth=None) -> BoundaryTides:
    """BoundaryTides with one slot per open-boundary face node: `constants(x, y)`
    gives a node's [n_c][6] C, S of eta, u, v; `mean(x, y)`, `source_depth(x, y)`
    its mean level and source depth (zeros by default)."""
    nodes = open_face_nodes(case)
    n1 = case.order + 1
    slot = np.full((case.n_elements, 4, n1), 0xFFFFFFFF, dtype=np.uint32)
    coef = np.zeros((len(nodes), len(omega), 6))
    me = np.zeros(len(nodes))
    sd = np.zeros(len(nodes))
    for i, (k, f, fi, x, y) in enumerate(nodes):
        slot[k, f, fi] = i
        coef[i] = constants(x, y)
        if mean is not None:
            me[i] = mean(x, y)
        if source_depth is not None:
            sd[i] = source_depth(x, y)
    return BoundaryTides(np.asarray(omega, dtype=np.float64), np.asarray(f_rate, dtype=np.float64), slot, coef, me,
                         sd, ramp, velocity, max_ratio)


# ---------------------------------------------------------------------------
# The nesting band (NestingRelaxation2D, src/boundary/ocean_nesting.rs:851-1060)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Nesting:
    """A relaxation band towards a parent model: per node `slot` [K*NN]
    (0xFFFFFFFF: not relaxed); the parent's snapshot `times` (Unix s) with
    `epoch` the Unix time of model time 0, `cubic` time interpolation; per slot
    `series` [n_times][3] (zeta, u, v; f32), band `weight` (1 on the open
    boundary) and parent `depth` (NaN: unknown); the timescale (s), the wet
    depth, momentum only, whether the parent has a velocity, the transport ratio
    limit (0: unscaled), the datum shift and the ramp (s)."""
    slot: np.ndarray
    times: np.ndarray
    series: np.ndarray
    weight: np.ndarray
    depth: np.ndarray
    timescale: float = 3600.0
    epoch: float = 0.0
    cubic: bool = False
    wet_depth: float = 0.05
    momentum_only: bool = False
    velocity: bool = False
    max_ratio: float = 0.0
    reference_level: float = 0.0
    ramp: float = 0.0


def cosine_profile(xi):
    """SpongeProfile::Cosine (source/boundary/sponge.rs:67-81)."""
    xi = min(max(xi, 0.0), 1.0)
    return 0.5 * (1.0 - math.cos(math.pi * xi))


def nesting_band(case, width, times, parent, parent_depth=None, profile=cosine_profile, **options) -> Nesting:
    """A nesting band over the open (TIDE) boundary: its face nodes (weight 1)
    and every node closer than `width` to one of its faces (weight
    profile(1 - d / width), as OceanModelState::new, ocean_nesting.rs:333-391,
    with d the distance to the nearest open face, straight between its corners);
    `parent(x, y, t)` gives the parent's (zeta, u, v) at a node and snapshot
    time, `parent_depth(x, y)` its depth."""
    x, y = node_coordinates(case.mesh, case.order)
    x, y = x.ravel(), y.ravel()
    nn = case.nn
    on = np.zeros(x.size, dtype=bool)
    for k, f, fi, _, _ in open_face_nodes(case):
        on[k * nn + face_node(case.order + 1, f, fi)] = True
    segs = []
    for k in range(case.n_elements):
        for f in range(4):
            if case.mesh.links[k, f] == BOUNDARY and case.mesh.tags[k, f] == TIDE:
                segs.append((case.mesh.corners[k, f], case.mesh.corners[k, (f + 1) % 4]))
    d = np.full(x.size, np.inf)
    for a, b in segs:
        ab = b - a
        t = np.clip(((x - a[0]) * ab[0] + (y - a[1]) * ab[1]) / (ab @ ab), 0.0, 1.0)
        d = np.minimum(d, np.hypot(x - (a[0] + t * ab[0]), y - (a[1] + t * ab[1])))
    times = np.asarray(times, dtype=np.float64)
    slot = np.full(x.size, 0xFFFFFFFF, dtype=np.uint32)
    series, weight, depth = [], [], []
    for n in range(x.size):
        if not (on[n] or d[n] < width):
            continue
        slot[n] = len(weight)
        weight.append(1.0 if on[n] else profile(1.0 - d[n] / width))
        depth.append(float("nan") if parent_depth is None else parent_depth(x[n], y[n]))
        series.append([parent(x[n], y[n], t) for t in times])
    return Nesting(slot, times, np.asarray(series, dtype=np.float32).reshape(len(weight), times.size, 3),
                   np.asarray(weight), np.asarray(depth), **options)


# ---------------------------------------------------------------------------
# Gridded weather (GriddedAtmosphere2D, src/source/swe_2d/gridded_atmosphere.rs)
# ---------------------------------------------------------------------------
DRAG_CONSTANT, DRAG_LARGE_POND, DRAG_WU, DRAG_SMITH, DRAG_YELLAND_TAYLOR = 0, 1, 2, 3, 4


@dataclasses.dataclass
class Weather:
    """Weather regridded to the nodes: snapshot `times` (Unix s) with `epoch`
    the Unix time of model time 0, `cubic` time interpolation; `series`
    [n_times][5][K*NN] f32 of u, v (m/s, mesh axes), dp/dx, dp/dy (Pa/m) and
    p - p_ref (Pa); the wind and pressure switches, the drag coefficient (a
    DRAG_* kind, `cd` for DRAG_CONSTANT), the densities, the depth below which a
    node is not forced and the ramp (s)."""
    times: np.ndarray
    series: np.ndarray
    epoch: float = 0.0
    cubic: bool = False
    wind: bool = True
    pressure: bool = True
    drag: int = DRAG_LARGE_POND
    cd: float = 0.0
    rho_air: float = 1.225
    rho_water: float = 1025.0
    h_min: float = 0.01
    ramp: float = 0.0


def weather_at_nodes(case, times, fields, **options) -> Weather:
    """Weather from `fields(x, y, t)` -> (u, v, dp/dx, dp/dy, p - p_ref), each an
    array over the nodes, at every snapshot time (as GriddedAtmosphere2D::
    regridded holds a weather model's fields at the nodes)."""
    x, y = node_coordinates(case.mesh, case.order)
    x, y = x.ravel(), y.ravel()
    times = np.asarray(times, dtype=np.float64)
    series = np.empty((times.size, 5, x.size), dtype=np.float32)
    for j, t in enumerate(times):
        for f, a in enumerate(fields(x, y, t)):
            series[j, f] = a
    return Weather(times, series, **options)


def boundary_tides_of(case, tide: Tide) -> BoundaryTides:
    """A uniform harmonic tide written as boundary tides (elevation only): per
    constituent C = f0 A cos(phi0), S = -f0 A sin(phi0) with phi0 = phase +
    (V0 + u0) degrees, omega' = omega + u_rate, f_rate / f0."""
    n = len(tide.amplitude)
    nod = tide.nodal if tide.nodal else [(1.0, 0.0, 0.0, 0.0, 0.0)] * n
    omega = [tide.omega[j] + math.radians(nod[j][4]) for j in range(n)]
    f_rate = [nod[j][1] / nod[j][0] for j in range(n)]
    c = np.zeros((n, 6))
    for j in range(n):
        f0, _, v0, u0, _ = nod[j]
        phi0 = tide.phase[j] + math.radians(v0 + u0)
        c[j, 0] = f0 * tide.amplitude[j] * math.cos(phi0)
        c[j, 1] = -f0 * tide.amplitude[j] * math.sin(phi0)
    return boundary_tides(case, omega, f_rate, lambda x, y: c, ramp=tide.ramp,
                          mean=(lambda x, y: tide.mean) if tide.mean else None)
