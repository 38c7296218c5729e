"""Shared small workloads and input perturbations for the parity tests."""
import numpy as np

from dgcom_amd import cases


def perturbed(case, seed=0, amp=0.05):
    """A seeded flow on the case's still water: depth noise where wet, random momenta."""
    rng = np.random.default_rng(seed)
    wet = case.h > 0.0
    h = case.h + amp * rng.standard_normal(case.h.size) * wet * np.minimum(case.h, 1.0)
    h = np.maximum(h, 0.0)
    hu = 0.3 * rng.standard_normal(h.size) * (h > 1e-3) * h
    hv = 0.3 * rng.standard_normal(h.size) * (h > 1e-3) * h
    return h, hu, hv


def small_unstructured(order=3):
    """The large unstructured workload's physics on a 24 x 24 jittered grid."""
    return cases.large_unstructured(n=24, length=2400.0, order=order, correlation=300.0, sd=20.0, mean_depth=15.0)


def periodic_distorted(order=2, n=6, length=600.0):
    """A doubly periodic, smoothly distorted mesh over a wavy bed, 10 m deep: every
    face interior, the periodic pairs among them."""
    tau = 2 * np.pi / length
    a = 0.04 * length

    def distort(x, y):
        sx, sy = np.sin(tau * x), np.sin(tau * y)
        return x + a * sx * sy + 0.5 * a * np.sin(tau * y), y - 0.7 * a * sx * sy + 0.4 * a * np.sin(tau * x)

    mesh = cases.structured_mesh(0.0, length, 0.0, length, n, n, periodic_x=True, periodic_y=True, distort=distort)
    x, y = cases.node_coordinates(mesh, order)
    bed = -(10.0 + 2.0 * np.sin(tau * x.ravel()) * np.cos(tau * y.ravel()))
    h = -bed
    return cases.Case("periodic", order, mesh, bed, np.full(mesh.n_elements, 0.025), 1.0e-4, cases.m2_tide(0.0),
                      h, 0.0 * h, 0.0 * h)


CASES = [
    ("estuary_n2", lambda: cases.tidal_estuary(nx=20, ny=4)),
    ("beach_n1", lambda: cases.drying_beach(order=1)),
    ("beach_n2_jitter", lambda: cases.drying_beach(order=2, jitter=0.2)),
    ("beach_n3_jitter", lambda: cases.drying_beach(order=3, jitter=0.2)),
    ("unstructured_n3", lambda: small_unstructured(3)),
    ("unstructured_n2", lambda: small_unstructured(2)),
    ("unstructured_n1", lambda: small_unstructured(1)),
    ("periodic_n3", lambda: periodic_distorted(3)),
]
