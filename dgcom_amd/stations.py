"""Stations: points where the model's water level is sampled, and the station
misfit whose gradient the adjoint computes.

A station is the reference's `Probe2D` (`src/solver/probe.rs:1-80`): its
element, the reference coordinates of the point in it (the inverse bilinear map
by Newton's method, `src/mesh/core/point_locator.rs:60-96`) and the element's
nodal Lagrange basis there (`DGOperators2D::interpolation_weights`,
`src/operators/operators_2d.rs:253-266`, node i = j (N+1) + a, r fastest), so
that the water level is eta = sum_i w_i (h_i + B_i).
"""
from __future__ import annotations

import numpy as np

from . import cases


def inverse_bilinear(corners, x, y):
    """(r, s) of the point (x, y) in the quad `corners` [4][2], or None."""
    (x0, y0), (x1, y1), (x2, y2), (x3, y3) = corners
    ax, ay = 0.25 * (x0 + x1 + x2 + x3), 0.25 * (y0 + y1 + y2 + y3)
    bx, by = 0.25 * (-x0 + x1 + x2 - x3), 0.25 * (-y0 + y1 + y2 - y3)
    cx, cy = 0.25 * (-x0 - x1 + x2 + x3), 0.25 * (-y0 - y1 + y2 + y3)
    dx, dy = 0.25 * (x0 - x1 + x2 - x3), 0.25 * (y0 - y1 + y2 - y3)
    noise = 16.0 * np.finfo(float).eps * max(abs(x), abs(y), abs(ax), abs(ay))
    r = s = 0.0
    for _ in range(50):
        fx = ax + bx * r + cx * s + dx * r * s - x
        fy = ay + by * r + cy * s + dy * r * s - y
        xr, xs = bx + dx * s, cx + dx * r
        yr, ys = by + dy * s, cy + dy * r
        det = xr * ys - xs * yr
        if det == 0.0 or not np.isfinite(det):
            return None
        dr = (ys * fx - xs * fy) / det
        ds = (-yr * fx + xr * fy) / det
        r -= dr
        s -= ds
        if max(abs(dr), abs(ds)) < 1e-14 * (1.0 + max(abs(r), abs(s))) + noise / np.sqrt(abs(det)):
            return r, s
        if max(abs(r), abs(s)) > 1e6:
            return None
    return None


def lagrange_basis(nodes, x):
    out = np.ones(len(nodes))
    for i, xi in enumerate(nodes):
        for j, xj in enumerate(nodes):
            if i != j:
                out[i] *= (x - xj) / (xi - xj)
    return out


def interpolation_weights(order, r, s):
    """Nodal basis values at (r, s), node i = j (N+1) + a."""
    x1 = cases.gll_nodes(order)
    lr, ls = lagrange_basis(x1, r), lagrange_basis(x1, s)
    return np.outer(ls, lr).ravel()


def locate(mesh, x, y, tol=1e-10):
    """(element, r, s) of the first element holding (x, y), or None."""
    c = mesh.corners
    xmin, xmax = c[:, :, 0].min(1), c[:, :, 0].max(1)
    ymin, ymax = c[:, :, 1].min(1), c[:, :, 1].max(1)
    for k in np.nonzero((xmin <= x) & (x <= xmax) & (ymin <= y) & (y <= ymax))[0]:
        rs = inverse_bilinear(c[k], x, y)
        if rs is not None and max(abs(rs[0]), abs(rs[1])) <= 1.0 + tol:
            return int(k), rs[0], rs[1]
    return None


class Stations:
    """Station elements and weights for a mesh and order."""

    def __init__(self, mesh, order, points):
        self.points = np.asarray(points, dtype=np.float64).reshape(-1, 2)
        el, w = [], []
        for x, y in self.points:
            loc = locate(mesh, x, y)
            if loc is None:
                raise ValueError(f"station ({x}, {y}) is outside the mesh")
            el.append(loc[0])
            w.append(interpolation_weights(order, loc[1], loc[2]))
        self.elements = np.array(el, dtype=np.int64)
        self.weights = np.array(w)
        self.nn = (order + 1) ** 2

    def eta(self, h, bed):
        """Water level at the stations from nodal depth and bed (host arrays; linear,
        so it also maps depth and bed tangents to the level's tangent)."""
        idx = self.elements[:, None] * self.nn + np.arange(self.nn)[None, :]
        return (self.weights * (h[idx] + bed[idx])).sum(1)


def random_stations(case, n, seed, min_depth=1.0):
    """n seeded points in elements whose every node is at least min_depth deep at rest."""
    rng = np.random.default_rng(seed)
    c = case.mesh.corners
    x0, x1 = c[:, :, 0].min(), c[:, :, 0].max()
    y0, y1 = c[:, :, 1].min(), c[:, :, 1].max()
    depth = -case.bed.reshape(case.n_elements, -1)
    pts = []
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 1000 * n:
            raise RuntimeError("could not place the stations in water")
        x, y = rng.uniform(x0, x1), rng.uniform(y0, y1)
        loc = locate(case.mesh, x, y)
        if loc is not None and depth[loc[0]].min() >= min_depth:
            pts.append((x, y))
    return Stations(case.mesh, case.order, pts)
