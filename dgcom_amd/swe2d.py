"""Host-side mirror of the reference's device interface for the 2D shallow-water
`WetDry` time step, over the C ABI of `include/dgswe.h` (ctypes).

The names follow the reference: `SWEPhysics2D` (`src/physics/builder.rs:83`),
whose `compute_rhs_into`, `compute_dt_ssp`, `element_dt`, `stage_where`,
`finish_where` and `post_process` are those of `PhysicsModule` /
`LocalTimeStepping` (`src/physics/traits.rs:61-163`,
`src/time/multirate.rs:296-391`); `DeviceState` is the device `Integrable`
(`gpu/src/swe2d.rs:2384-2396`); `Simulation.run` is the dt loop of
`Simulation::run_with_callback` (`src/simulation/runner.rs:264-483`).

There is no CPU fallback: without the built HIP library, or without a GPU,
constructing `SWEPhysics2D` raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)

SSP_RK3, SSP_RK43 = 0, 1
FRICTION, WET_DRY, SUBCELLS, INCOMING_WAVE, ENTROPY_STABLE = 1, 2, 4, 8, 16
BC_WALL, BC_TIDE = 0, 1

# Every entry point of include/dgswe.h
EXPORTS = (
    "dg_last_error", "dg_create", "dg_destroy", "dg_upload_form", "dg_set_tide", "dg_set_tide_nodal", "dg_state_alloc",
    "dg_state_free", "dg_state_upload", "dg_state_download", "dg_state_copy", "dg_rhs", "dg_face_fluxes", "dg_stage",
    "dg_element_dt", "dg_dt_ssp", "dg_step", "dg_run", "dg_is_finite", "dg_clip_count", "dg_masks",
    "dg_timing_enable", "dg_timing_read", "dg_sync", "dg_copy_bench",
    "dg_state_zero", "dg_state_wrap", "dg_set_stream", "dg_set_bed", "dg_set_manning", "dg_param_grads_zero",
    "dg_param_grads_read", "dg_rhs_vjp", "dg_stage_vjp", "dg_step_vjp", "dg_set_stations", "dg_sample_stations",
    "dg_tape_create", "dg_tape_destroy", "dg_misfit_forward", "dg_misfit_backward", "dg_tape_stats",
    "dg_tape_keep_stages", "dg_tape_records_used",
    "dg_revolve_count", "dg_batch_create", "dg_batch_destroy", "dg_batch_upload", "dg_batch_download", "dg_batch_run",
    "dg_batch_tuning", "dg_begin_coarse_step", "dg_stage_where", "dg_finish_where", "dg_depth_flags",
    "dg_multirate_create", "dg_multirate_destroy", "dg_multirate_assign_levels", "dg_multirate_step",
    "dg_multirate_levels", "dg_multirate_stats_read", "dg_multirate_run", "dg_fnv1a", "dg_set_boundary_tides",
    "dg_set_nesting", "dg_set_weather", "dg_set_inverse_barometer", "dg_set_nesting_boundary", "dg_masks_noalloc",
    "dg_set_manning_nodal", "dg_set_coriolis_beta", "dg_tune", "dg_launch_shape",
)
BATCH_GRAPH, BATCH_TUNE = 1, 2


class Form(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("n_elements", ctypes.c_int64), ("corners", _dp),
                ("face_links", ctypes.POINTER(ctypes.c_uint32)), ("face_tags", _u8p), ("bed", _dp),
                ("manning", _dp), ("manning_uniform", ctypes.c_double), ("coriolis_f0", ctypes.c_double),
                ("g", ctypes.c_double), ("h_min", ctypes.c_double), ("h_dry", ctypes.c_double),
                ("relaxation_time", ctypes.c_double), ("max_velocity", ctypes.c_double), ("flags", ctypes.c_int)]


class TideDesc(ctypes.Structure):
    _fields_ = [("mean", ctypes.c_double), ("ramp", ctypes.c_double), ("n", ctypes.c_int),
                ("amplitude", ctypes.c_double * 8), ("omega", ctypes.c_double * 8), ("phase", ctypes.c_double * 8)]


class StageDesc(ctypes.Structure):
    _fields_ = [("a", ctypes.c_double), ("b", ctypes.c_double), ("c", ctypes.c_double), ("time", ctypes.c_double)]


class StagePlanDesc(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("finest", ctypes.c_int), ("substep", ctypes.c_int),
                ("depth", ctypes.c_int), ("stage", ctypes.c_int), ("start", ctypes.c_double),
                ("input", ctypes.c_double), ("beta", ctypes.c_double), ("c", ctypes.c_double), ("last", ctypes.c_int)]


class MultirateStatsDesc(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("element_evaluations", ctypes.c_int64),
                ("global_evaluations", ctypes.c_int64), ("finest_level", ctypes.c_int),
                ("model_seconds", ctypes.c_double)]


_libs: dict = {}


def library_path(fma: bool = False) -> str:
    return os.path.join(HERE, "libdgswe_fma.so" if fma else "libdgswe.so")


def load(fma: bool = False):
    """The HIP library (raises if it was not built): the product build, or with
    `fma` the variant that contracts multiply-adds."""
    key = bool(fma)
    if key in _libs:
        return _libs[key]
    path = library_path(fma)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m dgcom_amd.build` (hipcc, gfx950)")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.dg_last_error.restype = ctypes.c_char_p
    L.dg_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.dg_destroy.argtypes = [vp]
    L.dg_upload_form.argtypes = [vp, ctypes.POINTER(Form)]
    L.dg_set_tide.argtypes = [vp, ctypes.POINTER(TideDesc)]
    L.dg_set_tide_nodal.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    L.dg_state_alloc.argtypes = [vp, ctypes.POINTER(vp)]
    L.dg_state_free.argtypes = [vp, vp]
    L.dg_state_upload.argtypes = [vp, vp, _dp, _dp, _dp]
    L.dg_state_download.argtypes = [vp, vp, _dp, _dp, _dp]
    L.dg_state_copy.argtypes = [vp, vp, vp]
    L.dg_rhs.argtypes = [vp, vp, ctypes.c_double, vp]
    L.dg_face_fluxes.argtypes = [vp, vp, ctypes.c_double, _dp]
    L.dg_stage.argtypes = [vp, vp, vp, vp, ctypes.POINTER(StageDesc)]
    L.dg_element_dt.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, _dp]
    L.dg_dt_ssp.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, _dp]
    L.dg_step.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.dg_run.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp]
    L.dg_is_finite.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_int)]
    L.dg_tune.argtypes = [vp, vp]
    L.dg_launch_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.dg_clip_count.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.dg_masks.argtypes = [vp, vp, _u8p, _u8p]
    L.dg_masks_noalloc.argtypes = [vp, vp, _u8p, _u8p]
    L.dg_timing_enable.argtypes = [vp, ctypes.c_int]
    L.dg_timing_read.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_int64), _dp, ctypes.POINTER(ctypes.c_int64)]
    L.dg_sync.argtypes = [vp]
    L.dg_copy_bench.argtypes = [vp, ctypes.c_size_t, ctypes.c_int, _dp]
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.dg_state_zero.argtypes = [vp, vp]
    L.dg_state_wrap.argtypes = [vp, vp, ctypes.POINTER(vp)]
    L.dg_set_stream.argtypes = [vp, vp]
    L.dg_set_bed.argtypes = [vp, vp, ctypes.c_int]
    L.dg_set_manning.argtypes = [vp, vp, ctypes.c_int]
    L.dg_set_manning_nodal.argtypes = [vp, vp, ctypes.c_int]
    L.dg_set_coriolis_beta.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.dg_param_grads_zero.argtypes = [vp]
    L.dg_param_grads_read.argtypes = [vp, vp, vp, ctypes.c_int]
    L.dg_rhs_vjp.argtypes = [vp, vp, ctypes.c_double, vp, vp]
    L.dg_stage_vjp.argtypes = [vp, vp, vp, ctypes.POINTER(StageDesc), vp, vp, vp]
    L.dg_step_vjp.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, vp, vp]
    L.dg_set_stations.argtypes = [vp, ctypes.c_int, i64p, _dp]
    L.dg_sample_stations.argtypes = [vp, vp, _dp]
    L.dg_tape_create.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.dg_tape_destroy.argtypes = [vp, vp]
    L.dg_misfit_forward.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                    _dp]
    L.dg_misfit_backward.argtypes = [vp, vp, ctypes.c_double, vp]
    L.dg_tape_stats.argtypes = [vp, i64p, i64p]
    L.dg_tape_keep_stages.argtypes = [vp, vp, ctypes.c_int]
    L.dg_tape_records_used.argtypes = [vp, i64p]
    L.dg_revolve_count.argtypes = [ctypes.c_int, ctypes.c_int, i64p]
    L.dg_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    L.dg_batch_destroy.argtypes = [vp]
    L.dg_batch_upload.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp]
    L.dg_batch_download.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp]
    L.dg_batch_run.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
    L.dg_batch_tuning.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.dg_begin_coarse_step.argtypes = [vp, _u8p, _u8p, ctypes.c_int]
    L.dg_stage_where.argtypes = [vp, vp, vp, ctypes.POINTER(StagePlanDesc), vp, vp]
    L.dg_finish_where.argtypes = [vp, vp, vp, ctypes.c_int]
    L.dg_depth_flags.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u8p]
    L.dg_set_boundary_tides.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int64, ctypes.POINTER(ctypes.c_uint32), _dp, _dp, _dp]
    L.dg_set_nesting.argtypes = ([vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int, _dp,
                                  ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_float), _dp, _dp]
                                 + [ctypes.c_double] * 2 + [ctypes.c_int] * 2 + [ctypes.c_double] * 3)
    L.dg_set_weather.argtypes = ([vp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
                                 + [ctypes.c_int] * 3 + [ctypes.c_double] * 5)
    L.dg_set_inverse_barometer.argtypes = [vp, ctypes.c_double]
    L.dg_set_nesting_boundary.argtypes = [vp, ctypes.c_int]
    L.dg_fnv1a.argtypes = [_u8p, ctypes.c_size_t, ctypes.c_uint64]
    L.dg_fnv1a.restype = ctypes.c_uint64
    L.dg_multirate_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.dg_multirate_destroy.argtypes = [vp]
    L.dg_multirate_assign_levels.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.dg_multirate_step.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double]
    L.dg_multirate_levels.argtypes = [vp, _u8p, _u8p, ctypes.POINTER(ctypes.c_int)]
    L.dg_multirate_stats_read.argtypes = [vp, ctypes.POINTER(MultirateStatsDesc)]
    L.dg_multirate_run.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, _dp, i64p]
    _libs[key] = L
    return L


def _check(L, status, what):
    if status != 0:
        raise RuntimeError(f"{what}: {L.dg_last_error().decode()} (status {status})")


def _p(a):
    return a.ctypes.data_as(_dp)


def revolve_count(n_steps: int, snapshots: int, fma: bool = False) -> int:
    """Forward steps (the first sweep included) of the checkpoint schedule for
    n_steps with `snapshots` stored states (host only, no device needed)."""
    L = load(fma)
    v = ctypes.c_int64()
    _check(L, L.dg_revolve_count(n_steps, snapshots, ctypes.byref(v)), "dg_revolve_count")
    return int(v.value)


class DeviceState:
    """A device-resident state (h, hu, hv), SoA FP64; with `device_ptr`, a view of
    caller memory (3*K*NN doubles, e.g. a torch tensor) that is not freed here."""

    def __init__(self, physics: "SWEPhysics2D", device_ptr=None):
        self.physics = physics
        self.handle = ctypes.c_void_p()
        if device_ptr is None:
            _check(physics.lib, physics.lib.dg_state_alloc(physics.ctx, ctypes.byref(self.handle)), "dg_state_alloc")
        else:
            _check(physics.lib, physics.lib.dg_state_wrap(physics.ctx, ctypes.c_void_p(device_ptr),
                                                          ctypes.byref(self.handle)), "dg_state_wrap")

    def zero(self):
        _check(self.physics.lib, self.physics.lib.dg_state_zero(self.physics.ctx, self.handle), "dg_state_zero")
        return self

    def upload(self, h, hu, hv):
        L = self.physics.lib
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (h, hu, hv)]
        n = self.physics.n_nodes_total
        if any(a.size != n for a in arrs):
            raise ValueError(f"state arrays must have {n} values each")
        _check(L, L.dg_state_upload(self.physics.ctx, self.handle, *(_p(a) for a in arrs)), "dg_state_upload")
        return self

    def download(self):
        L = self.physics.lib
        out = [np.empty(self.physics.n_nodes_total) for _ in range(3)]
        _check(L, L.dg_state_download(self.physics.ctx, self.handle, *(_p(a) for a in out)), "dg_state_download")
        return out

    def copy_from(self, other: "DeviceState"):
        L = self.physics.lib
        _check(L, L.dg_state_copy(self.physics.ctx, self.handle, other.handle), "dg_state_copy")

    def is_finite(self) -> bool:
        L = self.physics.lib
        f = ctypes.c_int()
        _check(L, L.dg_is_finite(self.physics.ctx, self.handle, ctypes.byref(f)), "dg_is_finite")
        return bool(f.value)

    def free(self):
        if self.handle and self.physics.ctx:
            self.physics.lib.dg_state_free(self.physics.ctx, self.handle)
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SWEPhysics2D:
    """The `WetDry` 2D shallow-water physics on one HIP device."""

    def __init__(self, order, mesh, bed, manning=None, manning_uniform=0.025, f0=0.0, tide=None, g=9.81,
                 h_min=1e-6, h_dry=1e-3, relaxation_time=1.0, max_velocity=20.0, friction=True, wet_dry=True,
                 subcells=True, device=0, fma=False, incoming_wave=False, entropy_stable=False):
        self.lib = L = load(fma)
        self.order = order
        self.nn = (order + 1) ** 2
        self.K = mesh.n_elements
        self.n_nodes_total = self.K * self.nn
        self.ctx = ctypes.c_void_p()
        _check(L, L.dg_create(device, ctypes.byref(self.ctx)), "dg_create")
        corners = np.ascontiguousarray(mesh.corners, dtype=np.float64)
        links = np.ascontiguousarray(mesh.links, dtype=np.uint32)
        tags = np.ascontiguousarray(mesh.tags, dtype=np.uint8)
        bed = np.ascontiguousarray(bed, dtype=np.float64)
        if bed.size != self.n_nodes_total:
            raise ValueError("bed must have K*(N+1)^2 values")
        man = None if manning is None else np.ascontiguousarray(manning, dtype=np.float64)
        # the EntropyStable faces take no subcells (swe_2d_split_form.rs:711-713 with its formulation)
        flags = (FRICTION if friction else 0) | (WET_DRY if wet_dry else 0) | \
                (SUBCELLS if subcells and not entropy_stable else 0) | (INCOMING_WAVE if incoming_wave else 0) | \
                (ENTROPY_STABLE if entropy_stable else 0)
        form = Form(order, self.K, _p(corners), links.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                    tags.ctypes.data_as(_u8p), _p(bed), None if man is None else _p(man), manning_uniform, f0, g,
                    h_min, h_dry, relaxation_time, max_velocity, flags)
        _check(L, L.dg_upload_form(self.ctx, ctypes.byref(form)), "dg_upload_form")
        self.flags, self.h_dry = flags, h_dry
        self.links = links
        if tide is not None:
            self.set_tide(tide)

    @classmethod
    def from_case(cls, case, **kw):
        return cls(case.order, case.mesh, case.bed, manning=case.manning, f0=case.f0, tide=case.tide, g=case.g,
                   h_min=case.h_min, h_dry=case.h_dry, incoming_wave=getattr(case, "incoming_wave", False), **kw)

    def set_tide(self, tide):
        n = len(tide.amplitude)
        if n > 8:
            raise ValueError("at most 8 tidal constituents")
        d = TideDesc()
        d.mean, d.ramp, d.n = tide.mean, tide.ramp, n
        for q in range(n):
            d.amplitude[q], d.omega[q], d.phase[q] = tide.amplitude[q], tide.omega[q], tide.phase[q]
        _check(self.lib, self.lib.dg_set_tide(self.ctx, ctypes.byref(d)), "dg_set_tide")
        nodal = getattr(tide, "nodal", ())
        if nodal:
            cols = [np.array([c[q] for c in nodal], dtype=np.float64) for q in range(5)]
            _check(self.lib, self.lib.dg_set_tide_nodal(self.ctx, *(_p(c) for c in cols)), "dg_set_tide_nodal")

    def set_boundary_tides(self, bt):
        """The open boundary's tides per face node (a `cases.BoundaryTides`), or None to clear them."""
        if bt is None:
            _check(self.lib, self.lib.dg_set_boundary_tides(self.ctx, 0, None, None, 0.0, 0, 0.0, 0, None, None,
                                                            None, None), "dg_set_boundary_tides")
            return
        om, fr, co, me, sd = (np.ascontiguousarray(a, dtype=np.float64) for a in (bt.omega, bt.f_rate, bt.coef,
                                                                                  bt.mean, bt.source_depth))
        sl = np.ascontiguousarray(bt.slot, dtype=np.uint32)
        _check(self.lib, self.lib.dg_set_boundary_tides(self.ctx, om.size, _p(om), _p(fr), bt.ramp, int(bt.velocity),
                                                        bt.max_ratio, me.size,
                                                        sl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _p(co),
                                                        _p(me), _p(sd)), "dg_set_boundary_tides")

    def set_nesting(self, ns):
        """A relaxation band towards a parent model (a `cases.Nesting`), or None to clear it."""
        if ns is None:
            _check(self.lib, self.lib.dg_set_nesting(self.ctx, 0, None, 0, None, 0.0, 0, None, None, None, 1.0, 0.05,
                                                     0, 0, 0.0, 0.0, 0.0), "dg_set_nesting")
            return
        sl = np.ascontiguousarray(ns.slot, dtype=np.uint32)
        ti = np.ascontiguousarray(ns.times, dtype=np.float64)
        se = np.ascontiguousarray(ns.series, dtype=np.float32)
        we = np.ascontiguousarray(ns.weight, dtype=np.float64)
        de = np.ascontiguousarray(ns.depth, dtype=np.float64)
        _check(self.lib, self.lib.dg_set_nesting(self.ctx, we.size, sl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                 ti.size, _p(ti), ns.epoch, int(ns.cubic),
                                                 se.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _p(we), _p(de),
                                                 ns.timescale, ns.wet_depth, int(ns.momentum_only),
                                                 int(ns.velocity), ns.max_ratio, ns.reference_level, ns.ramp),
               "dg_set_nesting")

    def set_weather(self, wx):
        """Gridded weather at the nodes (a `cases.Weather`), or None to clear it."""
        if wx is None:
            _check(self.lib, self.lib.dg_set_weather(self.ctx, 0, None, 0.0, 0, None, 0, 0, 1, 0.0, 1.225, 1025.0,
                                                     0.01, 0.0), "dg_set_weather")
            return
        ti = np.ascontiguousarray(wx.times, dtype=np.float64)
        se = np.ascontiguousarray(wx.series, dtype=np.float32)
        if se.shape != (ti.size, 5, self.n_nodes_total):
            raise ValueError(f"weather series of shape {se.shape}, expected {(ti.size, 5, self.n_nodes_total)}")
        _check(self.lib, self.lib.dg_set_weather(self.ctx, ti.size, _p(ti), wx.epoch, int(wx.cubic),
                                                 se.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(wx.wind),
                                                 int(wx.pressure), wx.drag, wx.cd, wx.rho_air, wx.rho_water, wx.h_min,
                                                 wx.ramp), "dg_set_weather")

    def set_nesting_boundary(self, on: bool = True):
        """The open boundary's external state taken from the nesting parent at its face nodes' slots."""
        _check(self.lib, self.lib.dg_set_nesting_boundary(self.ctx, int(on)), "dg_set_nesting_boundary")

    def set_inverse_barometer(self, factor: float):
        """Raise the open boundary's level by factor times the weather's inverse barometer (0: not)."""
        _check(self.lib, self.lib.dg_set_inverse_barometer(self.ctx, factor), "dg_set_inverse_barometer")

    def close(self):
        if self.ctx:
            self.lib.dg_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- states ---
    def state(self, h=None, hu=None, hv=None) -> DeviceState:
        s = DeviceState(self)
        if h is not None:
            s.upload(h, hu, hv)
        return s

    # --- PhysicsModule ---
    def compute_rhs_into(self, state: DeviceState, t: float, out: DeviceState):
        _check(self.lib, self.lib.dg_rhs(self.ctx, state.handle, t, out.handle), "dg_rhs")

    def face_fluxes(self, state: DeviceState, t: float) -> np.ndarray:
        """F* at every face node, (K, 4, N+1, 3), each element's own outward side."""
        out = np.empty(self.K * 4 * (self.order + 1) * 3)
        _check(self.lib, self.lib.dg_face_fluxes(self.ctx, state.handle, t, _p(out)), "dg_face_fluxes")
        return out.reshape(self.K, 4, self.order + 1, 3)

    def compute_dt_ssp(self, state: DeviceState, cfl: float, scheme: int = SSP_RK3) -> float:
        dt = ctypes.c_double()
        _check(self.lib, self.lib.dg_dt_ssp(self.ctx, state.handle, cfl, scheme, ctypes.byref(dt)), "dg_dt_ssp")
        return dt.value

    # --- LocalTimeStepping ---
    def element_dt(self, state: DeviceState, cfl: float, scheme: int = SSP_RK3) -> np.ndarray:
        out = np.empty(self.K)
        _check(self.lib, self.lib.dg_element_dt(self.ctx, state.handle, cfl, scheme, _p(out)), "dg_element_dt")
        return out

    def stage_where(self, base: DeviceState, input: DeviceState, out: DeviceState, a, b, c, time):
        d = StageDesc(a, b, c, time)
        _check(self.lib, self.lib.dg_stage(self.ctx, base.handle, input.handle, out.handle, ctypes.byref(d)),
               "dg_stage")

    def local_time_stepping(self) -> "LocalTimeStepping":
        """The element-subset stages of multirate stepping on this device
        (PhysicsModule::local_time_stepping, traits.rs:61-163)."""
        return LocalTimeStepping(self)

    def step(self, state: DeviceState, t: float, dt: float, scheme: int = SSP_RK3):
        _check(self.lib, self.lib.dg_step(self.ctx, state.handle, t, dt, scheme), "dg_step")

    def run(self, state: DeviceState, t: float, n_steps: int, cfl: float, scheme: int = SSP_RK3,
            record_dt: bool = False):
        te = ctypes.c_double()
        dts = np.empty(n_steps) if record_dt else None
        _check(self.lib, self.lib.dg_run(self.ctx, state.handle, t, n_steps, cfl, scheme, ctypes.byref(te),
                                         None if dts is None else _p(dts)), "dg_run")
        return (te.value, dts) if record_dt else te.value

    def tune(self, state: DeviceState) -> dict:
        """The stage's launch shape timed on `state` (dg_tune): kept for this context and for
        later forms of its order and size class in this process."""
        _check(self.lib, self.lib.dg_tune(self.ctx, state.handle), "dg_tune")
        return self.launch_shape()

    def launch_shape(self) -> dict:
        x, g = ctypes.c_int(), ctypes.c_int()
        _check(self.lib, self.lib.dg_launch_shape(self.ctx, ctypes.byref(x), ctypes.byref(g)), "dg_launch_shape")
        return {"xcd_mapping": bool(x.value), "subcell_groups": g.value}

    def negative_depth_clips(self) -> int:
        v = ctypes.c_uint64()
        _check(self.lib, self.lib.dg_clip_count(self.ctx, ctypes.byref(v)), "dg_clip_count")
        return int(v.value)

    def masks(self, state: DeviceState):
        nd = np.empty(self.n_nodes_total, dtype=np.uint8)
        es = np.empty(self.K, dtype=np.uint8)
        _check(self.lib, self.lib.dg_masks(self.ctx, state.handle, nd.ctypes.data_as(_u8p), es.ctypes.data_as(_u8p)),
               "dg_masks")
        return nd, es

    def masks_noalloc(self, state: DeviceState):
        """`masks` staged through the library's static device area (no allocation on the device)."""
        nd = np.empty(self.n_nodes_total, dtype=np.uint8)
        es = np.empty(self.K, dtype=np.uint8)
        _check(self.lib, self.lib.dg_masks_noalloc(self.ctx, state.handle, nd.ctypes.data_as(_u8p),
                                                   es.ctypes.data_as(_u8p)), "dg_masks_noalloc")
        return nd, es

    # --- measurement ---
    def timing(self, on: bool):
        _check(self.lib, self.lib.dg_timing_enable(self.ctx, int(on)), "dg_timing_enable")

    def timing_read(self):
        sm, dm = ctypes.c_double(), ctypes.c_double()
        sn, dn = ctypes.c_int64(), ctypes.c_int64()
        _check(self.lib, self.lib.dg_timing_read(self.ctx, ctypes.byref(sm), ctypes.byref(sn), ctypes.byref(dm),
                                                 ctypes.byref(dn)), "dg_timing_read")
        return sm.value, sn.value, dm.value, dn.value

    def sync(self):
        _check(self.lib, self.lib.dg_sync(self.ctx), "dg_sync")

    def copy_bench(self, nbytes: int, reps: int = 20) -> float:
        ms = ctypes.c_double()
        _check(self.lib, self.lib.dg_copy_bench(self.ctx, nbytes, reps, ctypes.byref(ms)), "dg_copy_bench")
        return ms.value

    # --- discrete adjoint ---
    def wrap(self, device_ptr: int) -> DeviceState:
        """A state over caller device memory of 3*K*NN doubles (h, hu, hv)."""
        return DeviceState(self, device_ptr)

    def set_stream(self, stream_handle):
        """Run on the caller's HIP stream (an integer handle), or the own stream (None)."""
        _check(self.lib, self.lib.dg_set_stream(self.ctx, ctypes.c_void_p(stream_handle or 0)), "dg_set_stream")

    def set_bed(self, bed=None, device_ptr=None):
        if device_ptr is not None:
            _check(self.lib, self.lib.dg_set_bed(self.ctx, ctypes.c_void_p(device_ptr), 1), "dg_set_bed")
            return
        b = np.ascontiguousarray(bed, dtype=np.float64)
        if b.size != self.n_nodes_total:
            raise ValueError("bed must have K*(N+1)^2 values")
        _check(self.lib, self.lib.dg_set_bed(self.ctx, b.ctypes.data, 0), "dg_set_bed")

    def set_manning_nodal(self, manning):
        """Manning n per node [K*NN] (SpatiallyVaryingManning2D); None: per element again."""
        if manning is None:
            _check(self.lib, self.lib.dg_set_manning_nodal(self.ctx, None, 0), "dg_set_manning_nodal")
            return
        a = np.ascontiguousarray(manning, dtype=np.float64)
        assert a.size == self.n_nodes_total
        _check(self.lib, self.lib.dg_set_manning_nodal(self.ctx, a.ctypes.data_as(ctypes.c_void_p), 0),
               "dg_set_manning_nodal")

    def set_coriolis_beta(self, beta, y0=0.0):
        """A beta-plane f = f0 + beta (y - y0) (CoriolisSource2D::beta_plane); beta = 0: the f-plane."""
        _check(self.lib, self.lib.dg_set_coriolis_beta(self.ctx, beta, y0), "dg_set_coriolis_beta")

    def set_manning(self, manning=None, device_ptr=None):
        if device_ptr is not None:
            _check(self.lib, self.lib.dg_set_manning(self.ctx, ctypes.c_void_p(device_ptr), 1), "dg_set_manning")
            return
        m = np.ascontiguousarray(manning, dtype=np.float64)
        if m.size != self.K:
            raise ValueError("manning must have K values")
        _check(self.lib, self.lib.dg_set_manning(self.ctx, m.ctypes.data, 0), "dg_set_manning")

    def param_grads_zero(self):
        _check(self.lib, self.lib.dg_param_grads_zero(self.ctx), "dg_param_grads_zero")

    def param_grads(self):
        """The accumulated (bed, Manning n) cotangents, on the host."""
        gb = np.empty(self.n_nodes_total)
        gm = np.empty(self.K)
        _check(self.lib, self.lib.dg_param_grads_read(self.ctx, gb.ctypes.data, gm.ctypes.data, 0),
               "dg_param_grads_read")
        return gb, gm

    def param_grads_to_device(self, bed_ptr: int, manning_ptr: int):
        _check(self.lib, self.lib.dg_param_grads_read(self.ctx, ctypes.c_void_p(bed_ptr), ctypes.c_void_p(manning_ptr),
                                                      1), "dg_param_grads_read")

    def rhs_vjp(self, state: DeviceState, t: float, cot_out: DeviceState, cot_in: DeviceState):
        _check(self.lib, self.lib.dg_rhs_vjp(self.ctx, state.handle, t, cot_out.handle, cot_in.handle), "dg_rhs_vjp")

    def stage_vjp(self, base: DeviceState, input: DeviceState, a, b, c, time, cot_out: DeviceState,
                  cot_base, cot_input: DeviceState):
        d = StageDesc(a, b, c, time)
        _check(self.lib, self.lib.dg_stage_vjp(self.ctx, base.handle, input.handle, ctypes.byref(d), cot_out.handle,
                                               None if cot_base is None else cot_base.handle, cot_input.handle),
               "dg_stage_vjp")

    def step_vjp(self, state: DeviceState, t: float, dt: float, cot_next: DeviceState, cot_state: DeviceState,
                 scheme: int = SSP_RK3):
        _check(self.lib, self.lib.dg_step_vjp(self.ctx, state.handle, t, dt, scheme, cot_next.handle,
                                              cot_state.handle), "dg_step_vjp")

    def set_stations(self, elements, weights):
        e = np.ascontiguousarray(elements, dtype=np.int64)
        w = np.ascontiguousarray(weights, dtype=np.float64)
        if w.size != e.size * self.nn:
            raise ValueError("weights must be [stations][(N+1)^2]")
        self.n_stations = e.size
        _check(self.lib, self.lib.dg_set_stations(self.ctx, e.size, e.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                  _p(w)), "dg_set_stations")

    def sample_stations(self, state: DeviceState) -> np.ndarray:
        out = np.empty(self.n_stations)
        _check(self.lib, self.lib.dg_sample_stations(self.ctx, state.handle, _p(out)), "dg_sample_stations")
        return out

    def tape(self, snapshots: int) -> "Tape":
        return Tape(self, snapshots)


class Tape:
    """Checkpoints of a station-misfit run for its reverse sweep (binomial schedule)."""

    def __init__(self, physics: SWEPhysics2D, snapshots: int):
        self.physics = physics
        self.snapshots = snapshots
        self.handle = ctypes.c_void_p()
        _check(physics.lib, physics.lib.dg_tape_create(physics.ctx, snapshots, ctypes.byref(self.handle)),
               "dg_tape_create")

    def forward(self, x0: DeviceState, t0: float, dts, every: int, obs, scheme: int = SSP_RK3) -> float:
        """J = sum (eta - obs)^2 over the samples after every `every`-th step; obs is
        [n_steps // every][stations]."""
        L = self.physics.lib
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        if obs.size != (dts.size // every) * self.physics.n_stations:
            raise ValueError("obs must be [n_steps // every][stations]")
        J = ctypes.c_double()
        _check(L, L.dg_misfit_forward(self.physics.ctx, self.handle, x0.handle, t0, dts.size, _p(dts), scheme, every,
                                      _p(obs), ctypes.byref(J)), "dg_misfit_forward")
        return J.value

    def backward(self, scale: float = 1.0, cot_x0: DeviceState = None):
        """Fills the parameter cotangents with scale * dJ/d(bed, n) (and cot_x0 with
        scale * dJ/dX0); returns (bed gradient, Manning gradient) on the host."""
        L = self.physics.lib
        _check(L, L.dg_misfit_backward(self.physics.ctx, self.handle, scale, None if cot_x0 is None else cot_x0.handle),
               "dg_misfit_backward")
        return self.physics.param_grads()

    def keep_stages(self, on: bool = True):
        """Keep step records (a step's stage values and RHS: 5 states per SSP-RK3
        step, snapshots + 1 records at most) from the next forward on, so that no
        reverse step recomputes its stages; the same gradient bit for bit."""
        _check(self.physics.lib, self.physics.lib.dg_tape_keep_stages(self.physics.ctx, self.handle, int(on)),
               "dg_tape_keep_stages")

    def records_used(self) -> int:
        """Reverse steps of the last backward that took their step's record."""
        u = ctypes.c_int64()
        _check(self.physics.lib, self.physics.lib.dg_tape_records_used(self.handle, ctypes.byref(u)),
               "dg_tape_records_used")
        return int(u.value)

    def stats(self):
        f, v = ctypes.c_int64(), ctypes.c_int64()
        _check(self.physics.lib, self.physics.lib.dg_tape_stats(self.handle, ctypes.byref(f), ctypes.byref(v)),
               "dg_tape_stats")
        return int(f.value), int(v.value)

    def close(self):
        if self.handle and self.physics.ctx:
            self.physics.lib.dg_tape_destroy(self.physics.ctx, self.handle)
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LocalTimeStepping:
    """`LocalTimeStepping` (src/time/multirate.rs:296-391) of a device physics,
    as `dgcom_amd.multirate.MultirateStepper` drives it: element time steps,
    the depth flags of `coarse_step_bound`, and the coarse step's passes through
    `dg_begin_coarse_step` / `dg_stage_where` / `dg_finish_where` (their element
    lists are the device's: those with a rate, or a level, of at least the
    pass's depth)."""

    def __init__(self, physics: SWEPhysics2D):
        self.physics = physics
        self.lib = physics.lib
        self.positivity_h_dry = physics.h_dry if physics.flags & WET_DRY else None
        # the line-local depth's h_dry (builder.rs:296-322: the WetDry kernel
        # changes at h_dry) and the node layout for the node-by-node rates
        self.line_local_h_dry = physics.h_dry
        self.n1 = physics.order + 1
        self.links = physics.links
        self.n_stages = 0

    def clone_state(self, x: DeviceState) -> DeviceState:
        s = DeviceState(self.physics)
        s.copy_from(x)
        return s

    def copy_state(self, dst: DeviceState, src: DeviceState):
        dst.copy_from(src)

    def free_state(self, x: DeviceState):
        x.free()

    def element_dt_state(self, state: DeviceState, cfl: float, scheme: int = SSP_RK3) -> np.ndarray:
        return self.physics.element_dt(state, cfl, scheme)

    def depth_flags(self, state: DeviceState, d0: float, d1: float, d2: float) -> np.ndarray:
        out = np.empty(self.physics.K, dtype=np.uint8)
        _check(self.lib, self.lib.dg_depth_flags(self.physics.ctx, state.handle, d0, d1, d2,
                                                 out.ctypes.data_as(_u8p)), "dg_depth_flags")
        return out

    def begin_coarse_step(self, level, rates):
        lv = np.ascontiguousarray(level, dtype=np.uint8)
        rt = np.ascontiguousarray(np.stack([np.asarray(r, dtype=np.uint8) for r in rates]))
        self.n_stages = rt.shape[0]
        _check(self.lib, self.lib.dg_begin_coarse_step(self.physics.ctx, lv.ctypes.data_as(_u8p),
                                                       rt.ctypes.data_as(_u8p), self.n_stages),
               "dg_begin_coarse_step")

    def stage_where(self, base: DeviceState, input: DeviceState, elements, plan, out: DeviceState, acc):
        if len(elements) == 0:
            return
        w = plan.weights
        d = StagePlanDesc(plan.t, plan.dt, plan.finest, plan.substep, plan.depth, plan.stage, w.start, w.input,
                          w.beta, w.c, int(plan.last))
        _check(self.lib, self.lib.dg_stage_where(self.physics.ctx, base.handle, input.handle, ctypes.byref(d),
                                                 out.handle, None if acc is None else acc.handle), "dg_stage_where")

    def finish_where(self, state: DeviceState, acc: DeviceState, elements, depth: int):
        if len(elements) == 0:
            return
        _check(self.lib, self.lib.dg_finish_where(self.physics.ctx, state.handle, acc.handle, depth),
               "dg_finish_where")


class Multirate:
    """Local time stepping with the library's own stepper (`dg_multirate_*`:
    MultirateStepper, src/time/multirate.rs:713-1330, in C++ beside the
    kernels): levels and stage rates from the device's element time steps, then
    the coarse step's passes, with no host round trip per pass."""

    LANDING_SLACK = 1e-9   # multirate.rs:588

    def __init__(self, physics: SWEPhysics2D, max_levels: int = 8, scheme: int = SSP_RK3, line_rates: bool = True):
        self.physics = physics
        self.lib = physics.lib
        self.scheme = scheme
        self.n_stages = 3 if scheme == SSP_RK3 else 4
        self.h = ctypes.c_void_p()
        _check(self.lib, self.lib.dg_multirate_create(physics.ctx, max_levels, scheme, int(line_rates),
                                                      ctypes.byref(self.h)), "dg_multirate_create")

    def assign_levels(self, state: DeviceState, cfl: float, dt_cap=None, landing=None):
        """(the coarse step fitted to `landing`, the step the elements allow)."""
        dt, stable = ctypes.c_double(), ctypes.c_double()
        _check(self.lib, self.lib.dg_multirate_assign_levels(self.h, state.handle, cfl, dt_cap or 0.0, landing or 0.0,
                                                             ctypes.byref(dt), ctypes.byref(stable)),
               "dg_multirate_assign_levels")
        return dt.value, stable.value

    def step(self, state: DeviceState, t: float, dt: float):
        _check(self.lib, self.lib.dg_multirate_step(self.h, state.handle, t, dt), "dg_multirate_step")

    def levels(self):
        K = self.physics.K
        level = np.zeros(K, dtype=np.uint8)
        rates = np.zeros((self.n_stages, K), dtype=np.uint8)
        finest = ctypes.c_int()
        _check(self.lib, self.lib.dg_multirate_levels(self.h, level.ctypes.data_as(_u8p), rates.ctypes.data_as(_u8p),
                                                      ctypes.byref(finest)), "dg_multirate_levels")
        return level, rates, finest.value

    def stats(self) -> dict:
        d = MultirateStatsDesc()
        _check(self.lib, self.lib.dg_multirate_stats_read(self.h, ctypes.byref(d)), "dg_multirate_stats_read")
        return {"steps": d.steps, "element_evaluations": d.element_evaluations,
                "global_evaluations": d.global_evaluations, "finest_level": d.finest_level,
                "model_seconds": d.model_seconds, "speedup": d.global_evaluations / max(d.element_evaluations, 1)}

    def run(self, state: DeviceState, t: float, t_end: float, cfl: float, dt_max=None, interval=None, callback=None):
        """Simulation::run_with_callback with a multirate integrator
        (runner.rs:298-455): coarse steps to t_end, landing on every output
        time t + k interval (callback(state, t) there). Returns (t, coarse steps)."""
        if callback is None:
            t_out, n = ctypes.c_double(), ctypes.c_int64()
            _check(self.lib, self.lib.dg_multirate_run(self.h, state.handle, t, t_end, cfl, dt_max or 0.0,
                                                       interval or 0.0, ctypes.byref(t_out), ctypes.byref(n)),
                   "dg_multirate_run")
            return t_out.value, n.value
        n, k, t_start = 0, 0, t
        while t < t_end:
            target = t_end if interval is None else min(t_start + (k + 1) * interval, t_end)
            fitted, dt = self.assign_levels(state, cfl, dt_max, target - t)
            if dt_max is not None:
                dt = min(dt, dt_max)
            dt = min(dt, fitted)
            landed = t + dt * (1.0 + self.LANDING_SLACK) >= target
            if landed:
                dt = target - t
            self.step(state, t, dt)
            t = target if landed else t + dt
            n += 1
            if interval is None:
                callback(state, t)
            elif t >= t_start + (k + 1) * interval:
                k += 1
                callback(state, t)
        return t, n

    def close(self):
        if self.h:
            self.lib.dg_multirate_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """The dt loop of `Simulation::run_with_callback` over a device physics
    (`src/simulation/runner.rs:264-483`): SSP-capped CFL, landing on `t_end`
    within LANDING_SLACK, a finiteness check per step."""

    LANDING_SLACK = 1e-9   # multirate.rs:588, the runner's (runner.rs:133)

    def __init__(self, physics: SWEPhysics2D, cfl: float, scheme: int = SSP_RK3, max_steps=None):
        self.physics = physics
        self.cfl = cfl
        self.scheme = scheme
        self.max_steps = max_steps

    def run(self, state: DeviceState, t: float, t_end: float, callback=None, interval=None):
        """Steps from t to t_end; with `interval`, every step that reaches the
        next callback time t_start + k interval lands on it and `callback` is
        called there (runner.rs:283-287, 405-455), else after every step."""
        n, k, t_start = 0, 0, t
        while t < t_end:
            if self.max_steps is not None and n >= self.max_steps:
                raise RuntimeError(f"Maximum step limit ({self.max_steps}) reached")
            target = t_end if interval is None else min(t_start + (k + 1) * interval, t_end)
            dt = self.physics.compute_dt_ssp(state, self.cfl, self.scheme)
            landed = t + dt * (1.0 + self.LANDING_SLACK) >= target
            if landed:
                dt = target - t
            self.physics.step(state, t, dt, self.scheme)
            t = target if landed else t + dt
            n += 1
            if not state.is_finite():
                raise RuntimeError(f"The state is not finite at t = {t:.1f} s: it has blown up")
            if interval is None:
                if callback is not None:
                    callback(state, t)
            elif t >= t_start + (k + 1) * interval:
                k += 1
                if callback is not None:
                    callback(state, t)
        return t, n


class Ensemble:
    """Independent runs advanced together (dg_batch): members of any order and size,
    each with its own adaptive time step and time; one launch per order per kernel
    for the whole ensemble and one host synchronisation per step."""

    def __init__(self, members, scheme: int = SSP_RK3):
        self.members = list(members)
        L = self.lib = self.members[0].lib
        arr = (ctypes.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        self.handle = ctypes.c_void_p()
        _check(L, L.dg_batch_create(len(self.members), arr, scheme, ctypes.byref(self.handle)), "dg_batch_create")
        self.t = np.zeros(len(self.members))

    def upload(self, m, h, hu, hv):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (h, hu, hv)]
        if any(a.size != self.members[m].n_nodes_total for a in arrs):
            raise ValueError("state arrays must have K*(N+1)^2 values each")
        _check(self.lib, self.lib.dg_batch_upload(self.handle, m, *(_p(a) for a in arrs)), "dg_batch_upload")

    def download(self, m):
        out = [np.empty(self.members[m].n_nodes_total) for _ in range(3)]
        _check(self.lib, self.lib.dg_batch_download(self.handle, m, *(_p(a) for a in out)), "dg_batch_download")
        return out

    def run(self, n_steps: int, cfl: float, graph: bool = True, tune: bool = False):
        opts = (BATCH_GRAPH if graph else 0) | (BATCH_TUNE if tune else 0)
        _check(self.lib, self.lib.dg_batch_run(self.handle, n_steps, cfl, opts, _p(self.t)), "dg_batch_run")
        return self.t.copy()

    def tuning(self, order: int):
        x, g = ctypes.c_int(), ctypes.c_int()
        _check(self.lib, self.lib.dg_batch_tuning(self.handle, order, ctypes.byref(x), ctypes.byref(g)),
               "dg_batch_tuning")
        return {"xcd_mapping": bool(x.value), "subcell_groups": g.value}

    def close(self):
        if self.handle:
            self.lib.dg_batch_destroy(self.handle)
        self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
