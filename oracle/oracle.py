"""ORACLE — test infrastructure only.

ctypes access to the CPU restatement of the reference's 2D shallow-water path
(`liboracle_swe2d.so`, built from this folder by its Makefile). Only the tests,
the smoke check and the benchmark's CPU-baseline leg use it, as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_swe2d.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_lib = None

FLAG_SUBCELLS, FLAG_FRICTION, FLAG_WET_DRY, FLAG_MANUFACTURED, FLAG_INCOMING = 1, 2, 4, 8, 16
SURFACE_HLL, SURFACE_ES, SURFACE_EC = 0, 1, 2


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def mutant(name: str):
    """A planted-mistake build of the restatement (NO_BED, LIFT, TRANSPOSE), bound
    like lib(); use with `using(...)`."""
    path = os.path.join(HERE, f"liboracle_swe2d_mutant_{name.lower()}.so")
    subprocess.run(["make", "-s", "-C", HERE, "mutant", f"MUTANT={name}"], check=True)
    return _bind(ctypes.CDLL(path))


class using:
    """Context manager: route Oracle objects created inside to another build."""

    def __init__(self, L):
        self.L = L

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = self.L
        return self.L

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved


def _bind(L):
    """Declare the C signatures on a loaded build."""
    if True:
        L.oc_create.restype = ctypes.c_void_p
        L.oc_create.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _u32p, _u8p, _dp, _dp, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.oc_last_error.restype = ctypes.c_char_p
        L.oc_destroy.argtypes = [ctypes.c_void_p]
        L.oc_set_tide.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
        L.oc_set_tide_nodal.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oc_tide_elevation.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oc_tide_elevation.restype = ctypes.c_double
        L.oc_rhs.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp]
        L.oc_face_mass.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.oc_stage.argtypes = [ctypes.c_void_p] + [_dp] * 9 + [ctypes.c_double] * 4 + [ctypes.c_int]
        L.oc_step.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.oc_dt.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int]
        L.oc_dt.restype = ctypes.c_double
        L.oc_element_dt.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.oc_run.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                             ctypes.c_int, _dp]
        L.oc_run.restype = ctypes.c_double
        L.oc_run_timed.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_int, _dp]
        L.oc_run_timed.restype = ctypes.c_double
        L.oc_clips.argtypes = [ctypes.c_void_p]
        L.oc_clips.restype = ctypes.c_uint64
        L.oc_masks.argtypes = [ctypes.c_void_p, _dp, _u8p, _u8p]
        L.oc_operators.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.oc_geometry.argtypes = [ctypes.c_void_p] + [_dp] * 6
        L.oc_cbrt.argtypes = [ctypes.c_double]
        L.oc_cbrt.restype = ctypes.c_double
        L.oc_set_threads.argtypes = [ctypes.c_int]
        L.oc_max_threads.restype = ctypes.c_int
        L.oc_face_mass.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.oc_tlm_create.restype = ctypes.c_void_p
        L.oc_tlm_create.argtypes = L.oc_create.argtypes
        L.oc_tlm_destroy.argtypes = [ctypes.c_void_p]
        L.oc_tlm_set_tide.argtypes = L.oc_set_tide.argtypes
        L.oc_tlm_set_tide_nodal.argtypes = L.oc_set_tide_nodal.argtypes
        L.oc_tlm_set_param_tangents.argtypes = [ctypes.c_void_p, _dp, _dp]
        _pp = ctypes.POINTER(_dp)
        L.oc_tlm_stage.argtypes = [ctypes.c_void_p, _pp, _pp, _pp, _pp] + [ctypes.c_double] * 4 + [ctypes.c_int, _pp, _pp]
        L.oc_tlm_rhs.argtypes = [ctypes.c_void_p, _pp, _pp, ctypes.c_double, _pp, _pp]
        L.oc_tlm_run.argtypes = [ctypes.c_void_p, _pp, _pp, ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int]
        L.oc_tlm_run.restype = ctypes.c_double
        L.oc_run_dts.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int]
        L.oc_run_dts.restype = ctypes.c_double
        L.oc_stage_where.argtypes = ([ctypes.c_void_p] + [_dp] * 6 + [_u32p, ctypes.c_int64, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
                                     + [ctypes.c_double] * 4 + [ctypes.c_int, _u8p, _u8p] + [_dp] * 6)
        L.oc_finish_where.argtypes = [ctypes.c_void_p] + [_dp] * 6 + [_u32p, ctypes.c_int64]
        L.oc_set_boundary_tides.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_int64, _u32p, _dp, _dp, _dp]
        L.oc_set_nesting.argtypes = ([ctypes.c_void_p, ctypes.c_int64, _u32p, ctypes.c_int, _dp, ctypes.c_double,
                                      ctypes.c_int, ctypes.POINTER(ctypes.c_float), _dp, _dp]
                                     + [ctypes.c_double] * 2 + [ctypes.c_int] * 2 + [ctypes.c_double] * 3)
        L.oc_set_weather.argtypes = ([ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_float)] + [ctypes.c_int] * 3 + [ctypes.c_double] * 5)
        L.oc_set_inverse_barometer.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oc_set_nesting_boundary.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oc_set_relaxation_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oc_set_beta.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.oc_set_manning_nodal.argtypes = [ctypes.c_void_p, _dp]
        L.oc_hypot.argtypes = [ctypes.c_double, ctypes.c_double]
        L.oc_hypot.restype = ctypes.c_double
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def hypot(x, y):
    """The reference's hypot (math.rs:205-249) as the restatement computes it."""
    return lib().oc_hypot(x, y)


def operators(order):
    n1 = order + 1
    x = np.zeros(n1)
    w = np.zeros(n1)
    d = np.zeros(n1 * n1)
    lib().oc_operators(order, _p(x), _p(w), _p(d))
    return x, w, d.reshape(n1, n1)


class Oracle:
    """The CPU restatement on one problem (a `dgcom_amd.cases.Case`-shaped input)."""

    def __init__(self, order, mesh, bed, manning=None, f0=0.0, g=9.81, h_min=1e-6, h_dry=1e-3, surface=SURFACE_HLL,
                 subcells=True, friction=True, wet_dry=True, manufactured=False, tide=None, incoming_wave=False):
        self._L = L = lib()
        self.order = order
        self.nn = (order + 1) ** 2
        self.K = mesh.n_elements
        self._corners = np.ascontiguousarray(mesh.corners, dtype=np.float64)
        self._links = np.ascontiguousarray(mesh.links, dtype=np.uint32)
        self._tags = np.ascontiguousarray(mesh.tags, dtype=np.uint8)
        self._bed = np.ascontiguousarray(bed, dtype=np.float64)
        man = None if manning is None else np.ascontiguousarray(manning, dtype=np.float64)
        flags = (FLAG_SUBCELLS if subcells else 0) | (FLAG_FRICTION if friction else 0) | \
                (FLAG_WET_DRY if wet_dry else 0) | (FLAG_MANUFACTURED if manufactured else 0) | \
                (FLAG_INCOMING if incoming_wave else 0)
        self.h = L.oc_create(order, self.K, _p(self._corners), self._links.ctypes.data_as(_u32p),
                             self._tags.ctypes.data_as(_u8p), _p(self._bed), None if man is None else _p(man),
                             g, h_min, h_dry, f0, surface, flags)
        if not self.h:
            raise RuntimeError(L.oc_last_error().decode())
        self.positivity_h_dry = h_dry if wet_dry else None
        self.line_local_h_dry = h_dry
        self.n1 = order + 1
        self.links = self._links
        if tide is not None:
            self.set_tide(tide)

    @classmethod
    def from_case(cls, case, **kw):
        return cls(case.order, case.mesh, case.bed, manning=case.manning, f0=case.f0, g=case.g, h_min=case.h_min,
                   h_dry=case.h_dry, tide=case.tide, incoming_wave=getattr(case, "incoming_wave", False), **kw)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.oc_destroy(self.h)
            self.h = None

    def set_tide(self, tide):
        n = len(tide.amplitude)
        a = np.array(tide.amplitude, dtype=np.float64)
        w = np.array(tide.omega, dtype=np.float64)
        p = np.array(tide.phase, dtype=np.float64)
        self._L.oc_set_tide(self.h, tide.mean, tide.ramp, n, _p(a), _p(w), _p(p))
        nodal = getattr(tide, "nodal", ())
        if nodal:
            cols = [np.array([c[q] for c in nodal], dtype=np.float64) for q in range(5)]
            self._L.oc_set_tide_nodal(self.h, n, *(_p(c) for c in cols))

    def set_boundary_tides(self, bt):
        """BoundaryTides per boundary face node (a `cases.BoundaryTides`)."""
        self._bt = [np.ascontiguousarray(a, dtype=np.float64) for a in (bt.omega, bt.f_rate, bt.coef, bt.mean,
                                                                        bt.source_depth)]
        self._bt_slot = np.ascontiguousarray(bt.slot, dtype=np.uint32)
        om, fr, co, me, sd = self._bt
        self._L.oc_set_boundary_tides(self.h, len(om), _p(om), _p(fr), bt.ramp, int(bt.velocity), bt.max_ratio,
                                      me.size, self._bt_slot.ctypes.data_as(_u32p), _p(co), _p(me), _p(sd))

    def set_nesting(self, ns):
        """The relaxation band towards a parent (a `cases.Nesting`)."""
        self._ns = (np.ascontiguousarray(ns.slot, dtype=np.uint32), np.ascontiguousarray(ns.times, dtype=np.float64),
                    np.ascontiguousarray(ns.series, dtype=np.float32), np.ascontiguousarray(ns.weight, dtype=np.float64),
                    np.ascontiguousarray(ns.depth, dtype=np.float64))
        sl, ti, se, we, de = self._ns
        self._L.oc_set_nesting(self.h, we.size, sl.ctypes.data_as(_u32p), ti.size, _p(ti), ns.epoch, int(ns.cubic),
                               se.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _p(we), _p(de), ns.timescale,
                               ns.wet_depth, int(ns.momentum_only), int(ns.velocity), ns.max_ratio,
                               ns.reference_level, ns.ramp)

    def set_weather(self, wx):
        """Gridded weather at the nodes (a `cases.Weather`)."""
        self._wx = (np.ascontiguousarray(wx.times, dtype=np.float64), np.ascontiguousarray(wx.series, dtype=np.float32))
        ti, se = self._wx
        self._L.oc_set_weather(self.h, ti.size, _p(ti), wx.epoch, int(wx.cubic),
                               se.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(wx.wind), int(wx.pressure),
                               wx.drag, wx.cd, wx.rho_air, wx.rho_water, wx.h_min, wx.ramp)

    def set_inverse_barometer(self, factor):
        """The open boundary's level raised by factor times the weather's inverse barometer."""
        self._L.oc_set_inverse_barometer(self.h, factor)

    def set_nesting_boundary(self, on=True):
        """The open boundary's external state taken from the nesting parent at its face nodes' slots."""
        self._L.oc_set_nesting_boundary(self.h, int(on))

    def set_relaxation_time(self, tau):
        """The thin-layer relaxation time (wetting_drying.rs:82-89)."""
        self._L.oc_set_relaxation_time(self.h, tau)

    def set_beta(self, beta, y0=0.0):
        """A beta-plane f = f0 + beta (y - y0) (coriolis.rs:101-121)."""
        self._L.oc_set_beta(self.h, beta, y0)

    def set_manning_nodal(self, n):
        """Manning n per node [K*NN] (SpatiallyVaryingManning2D); None: per element."""
        self._mn_node = None if n is None else np.ascontiguousarray(n, dtype=np.float64)
        self._L.oc_set_manning_nodal(self.h, None if n is None else _p(self._mn_node))

    def tide_elevation(self, t):
        return self._L.oc_tide_elevation(self.h, t)

    def rhs(self, h, hu, hv, t=0.0):
        h, hu, hv = (np.ascontiguousarray(a, dtype=np.float64) for a in (h, hu, hv))
        out = [np.empty_like(h) for _ in range(3)]
        self._L.oc_rhs(self.h, _p(h), _p(hu), _p(hv), t, *(_p(o) for o in out))
        return out

    def face_mass(self, h, hu, hv, t=0.0):
        out = np.empty(self.K * 4 * (self.order + 1))
        self._L.oc_face_mass(self.h, _p(h), _p(hu), _p(hv), t, _p(out))
        return out.reshape(self.K, 4, self.order + 1)

    def stage(self, base, inp, a, b, c, time, last):
        out = [np.empty_like(base[0]) for _ in range(3)]
        self._L.oc_stage(self.h, *(_p(x) for x in base), *(_p(x) for x in inp), *(_p(o) for o in out), a, b, c, time,
                       int(last))
        return out

    # --- LocalTimeStepping (multirate.rs:296-391) over host states [h, hu, hv] ---
    def clone_state(self, x):
        return [np.array(a, dtype=np.float64, copy=True) for a in x]

    def copy_state(self, dst, src):
        for d, s_ in zip(dst, src):
            d[...] = s_

    def begin_coarse_step(self, level, rates):
        self._level = np.ascontiguousarray(level, dtype=np.uint8)
        self._rates = [np.ascontiguousarray(r, dtype=np.uint8) for r in rates]

    def stage_where(self, base, inp, elements, plan, out, acc):
        """builder.rs:652-737 on the listed elements; element k takes plan's
        element_stage of its level and (last stage) rate."""
        el = np.ascontiguousarray(elements, dtype=np.uint32)
        if el.size == 0:
            return
        w = plan.weights
        a3 = acc if acc is not None else [None, None, None]
        self._L.oc_stage_where(self.h, *(_p(x) for x in base), *(_p(x) for x in inp), el.ctypes.data_as(_u32p),
                               el.size, plan.t, plan.dt, plan.finest, plan.substep, plan.depth, w.start, w.input,
                               w.beta, w.c, int(plan.last), self._level.ctypes.data_as(_u8p),
                               self._rates[plan.stage].ctypes.data_as(_u8p), *(_p(o) for o in out),
                               *((_p(x) if x is not None else None) for x in a3))

    def finish_where(self, state, acc, elements, depth):
        """builder.rs:739-782 on the listed elements."""
        el = np.ascontiguousarray(elements, dtype=np.uint32)
        if el.size == 0:
            return
        self._L.oc_finish_where(self.h, *(_p(x) for x in state), *(_p(x) for x in acc), el.ctypes.data_as(_u32p),
                                el.size)

    def element_dt_state(self, x, cfl, scheme=0):
        return self.element_dt(x[0], x[1], x[2], cfl, scheme)

    def depth_flags(self, x, d0, d1, d2):
        h = np.asarray(x[0]).reshape(self.K, -1)
        return ((h > d0).all(1).astype(np.uint8) | ((h > d1).all(1).astype(np.uint8) << 1)
                | ((h > d2).all(1).astype(np.uint8) << 2))

    def step(self, h, hu, hv, t, dt, scheme=0):
        """One SSP step in place."""
        self._L.oc_step(self.h, _p(h), _p(hu), _p(hv), t, dt, scheme)

    def dt(self, h, hu, hv, cfl, scheme=0):
        return self._L.oc_dt(self.h, _p(h), _p(hu), _p(hv), cfl, scheme)

    def element_dt(self, h, hu, hv, cfl, scheme=0):
        out = np.empty(self.K)
        self._L.oc_element_dt(self.h, _p(h), _p(hu), _p(hv), cfl, scheme, _p(out))
        return out

    def run(self, h, hu, hv, t, n_steps, cfl, scheme=0, record_dt=False):
        """n_steps adaptive SSP steps in place; returns the end time (and the dts)."""
        dts = np.empty(n_steps) if record_dt else None
        te = self._L.oc_run(self.h, _p(h), _p(hu), _p(hv), t, n_steps, cfl, scheme, None if dts is None else _p(dts))
        return (te, dts) if record_dt else te

    def run_dts(self, h, hu, hv, t, dts, scheme=0):
        """Steps at the prescribed dts, in place."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        return self._L.oc_run_dts(self.h, _p(h), _p(hu), _p(hv), t, len(dts), _p(dts), scheme)

    def run_timed(self, h, hu, hv, t, n_steps, cfl, scheme=0):
        sec = ctypes.c_double()
        te = self._L.oc_run_timed(self.h, _p(h), _p(hu), _p(hv), t, n_steps, cfl, scheme, ctypes.byref(sec))
        return te, sec.value

    def clips(self):
        return int(self._L.oc_clips(self.h))

    def masks(self, h):
        nd = np.empty(self.K * self.nn, dtype=np.uint8)
        es = np.empty(self.K, dtype=np.uint8)
        self._L.oc_masks(self.h, _p(np.ascontiguousarray(h)), nd.ctypes.data_as(_u8p), es.ctypes.data_as(_u8p))
        return nd, es

    def geometry(self):
        out = [np.empty(self.K * self.nn) for _ in range(6)]
        self._L.oc_geometry(self.h, *(_p(o) for o in out))
        return dict(zip(("det_j", "mass", "jrx", "jry", "jsx", "jsy"), out))


def _ptrs(arrs):
    """A C array of double pointers to three contiguous float64 arrays."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
    return (_dp * 3)(*(_p(a) for a in arrs)), arrs


class TangentLinear:
    """The restatement on dual numbers: directional derivatives of stages and runs
    with respect to the states, the nodal bed and per-element Manning n (time steps
    prescribed, as a discrete adjoint freezes them)."""

    def __init__(self, order, mesh, bed, manning=None, f0=0.0, g=9.81, h_min=1e-6, h_dry=1e-3, tide=None,
                 subcells=True, friction=True, wet_dry=True, incoming_wave=False):
        self._L = L = lib()
        self.K = mesh.n_elements
        self.nn = (order + 1) ** 2
        self._keep = [np.ascontiguousarray(mesh.corners, dtype=np.float64),
                      np.ascontiguousarray(mesh.links, dtype=np.uint32),
                      np.ascontiguousarray(mesh.tags, dtype=np.uint8),
                      np.ascontiguousarray(bed, dtype=np.float64),
                      None if manning is None else np.ascontiguousarray(manning, dtype=np.float64)]
        c, l, t, b, m = self._keep
        flags = (FLAG_SUBCELLS if subcells else 0) | (FLAG_FRICTION if friction else 0) | \
                (FLAG_WET_DRY if wet_dry else 0) | (FLAG_INCOMING if incoming_wave else 0)
        self.h = L.oc_tlm_create(order, self.K, _p(c), l.ctypes.data_as(_u32p), t.ctypes.data_as(_u8p), _p(b),
                                 None if m is None else _p(m), g, h_min, h_dry, f0, SURFACE_HLL, flags)
        if not self.h:
            raise RuntimeError(L.oc_last_error().decode())
        if tide is not None:
            a = np.array(tide.amplitude, dtype=np.float64)
            w = np.array(tide.omega, dtype=np.float64)
            ph = np.array(tide.phase, dtype=np.float64)
            L.oc_tlm_set_tide(self.h, tide.mean, tide.ramp, len(a), _p(a), _p(w), _p(ph))
            nodal = getattr(tide, "nodal", ())
            if nodal:
                cols = [np.array([c[q] for c in nodal], dtype=np.float64) for q in range(5)]
                L.oc_tlm_set_tide_nodal(self.h, len(a), *(_p(c) for c in cols))

    @classmethod
    def from_case(cls, case, **kw):
        return cls(case.order, case.mesh, case.bed, manning=case.manning, f0=case.f0, g=case.g, h_min=case.h_min,
                   h_dry=case.h_dry, tide=case.tide, incoming_wave=getattr(case, "incoming_wave", False), **kw)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.oc_tlm_destroy(self.h)
            self.h = None

    def set_param_tangents(self, dbed=None, dmanning=None):
        self._dp_keep = [None if dbed is None else np.ascontiguousarray(dbed, dtype=np.float64),
                         None if dmanning is None else np.ascontiguousarray(dmanning, dtype=np.float64)]
        a, b = self._dp_keep
        self._L.oc_tlm_set_param_tangents(self.h, None if a is None else _p(a), None if b is None else _p(b))

    def stage(self, base, dbase, inp, dinp, a, b, c, time, last):
        pb, kb = _ptrs(base)
        pdb, kdb = _ptrs(dbase)
        pi, ki = _ptrs(inp)
        pdi, kdi = _ptrs(dinp)
        out = [np.empty(self.K * self.nn) for _ in range(3)]
        dout = [np.empty(self.K * self.nn) for _ in range(3)]
        po, _ = _ptrs(out)
        pdo, _ = _ptrs(dout)
        self._L.oc_tlm_stage(self.h, pb, pdb, pi, pdi, a, b, c, time, int(last), po, pdo)
        return out, dout

    def rhs(self, inp, dinp, time):
        pi, ki = _ptrs(inp)
        pdi, kdi = _ptrs(dinp)
        out = [np.empty(self.K * self.nn) for _ in range(3)]
        dout = [np.empty(self.K * self.nn) for _ in range(3)]
        po, _ = _ptrs(out)
        pdo, _ = _ptrs(dout)
        self._L.oc_tlm_rhs(self.h, pi, pdi, time, po, pdo)
        return out, dout

    def run(self, x, dx, t, dts, scheme=0):
        """Steps at the prescribed dts, in place on the value and tangent arrays."""
        px, kx = _ptrs(x)
        pdx, kdx = _ptrs(dx)
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        te = self._L.oc_tlm_run(self.h, px, pdx, t, len(dts), _p(dts), scheme)
        for a, b in zip(x, kx):
            a[...] = b
        for a, b in zip(dx, kdx):
            a[...] = b
        return te


def set_threads(n):
    lib().oc_set_threads(n)


def max_threads():
    return lib().oc_max_threads()


def cbrt(x):
    return lib().oc_cbrt(x)
