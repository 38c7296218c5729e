"""ctypes access to the CPU emulation of the product's HIP kernels (tests only)."""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")
_dp = ctypes.POINTER(ctypes.c_double)
_lib = None
M_LEVELS = 8   # DG_MAX_LEVELS of include/dgswe.h


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = ctypes.CDLL(os.path.join(HERE, "libswe_emu.so"))
        L.emu_stage.restype = ctypes.c_ulonglong
        L.emu_stage.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                _dp, _dp, _dp, _dp, _dp] + [ctypes.c_double] * 10 + [ctypes.c_int, ctypes.c_int, _dp]
        L.emu_dt.restype = ctypes.c_double
        L.emu_dt.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, _dp, ctypes.c_void_p, ctypes.c_void_p] + \
            [ctypes.c_double] * 7 + [ctypes.c_int, _dp]
        L.emu_stage_vjp.restype = None
        L.emu_stage_vjp.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp] + \
            [ctypes.c_double] * 10 + [ctypes.c_int, ctypes.c_int]
        vp = ctypes.c_void_p
        L.emu_stage_where.restype = ctypes.c_ulonglong
        L.emu_stage_where.argtypes = ([ctypes.c_int, ctypes.c_int64, _dp, vp, vp, vp, _dp, _dp, _dp, _dp, _dp, _dp,
                                       vp, ctypes.c_int64, vp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       _dp, _dp] + [ctypes.c_double] * 6 + [ctypes.c_int])
        L.emu_stage_where_forced.restype = ctypes.c_ulonglong
        L.emu_stage_where_forced.argtypes = ([ctypes.c_int, ctypes.c_int64, _dp, vp, vp, vp, _dp, _dp, _dp, _dp, _dp,
                                              _dp, vp, ctypes.c_int64, vp, vp, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, _dp, _dp] + [ctypes.c_double] * 6 +
                                             [ctypes.c_int, vp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, _dp, _dp, _dp])
        L.emu_coriolis_nodes.restype = None
        L.emu_coriolis_nodes.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _dp]
        L.emu_lvl_moment.restype = ctypes.c_int
        L.emu_finish_where.restype = ctypes.c_ulonglong
        L.emu_finish_where.argtypes = [ctypes.c_int, ctypes.c_int64, _dp, vp, _dp, _dp, vp, ctypes.c_int64,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int]
        for fn in (L.emu_multirate_assign, L.emu_multirate_assign_device):
            fn.restype = None
            fn.argtypes = [ctypes.c_int64, ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, _dp, vp, ctypes.c_int, _dp, _dp, _dp, _dp, vp, vp, vp]
        L.emu_init()
        _lib = L
    return _lib


def parallelogram(c):
    v = c.reshape(-1, 8)
    bx = v[:, 0] - v[:, 2] + v[:, 4] - v[:, 6]
    by = v[:, 1] - v[:, 3] + v[:, 5] - v[:, 7]
    scale = np.zeros(len(v))
    for i in range(4):
        j = (i + 1) % 4
        scale = np.fmax(scale, np.hypot(v[:, 2 * j] - v[:, 2 * i], v[:, 2 * j + 1] - v[:, 2 * i + 1]))
    return (np.hypot(bx, by) <= 1.0e-12 * scale).astype(np.uint8)


class EmulatedKernels:
    """The product's stage and time-step kernels run on host threads."""

    def __init__(self, case, flags=None):
        self.case = case
        self.order = case.order
        self.K = case.n_elements
        self.corners = np.ascontiguousarray(case.mesh.corners, dtype=np.float64).ravel()
        self.links = np.ascontiguousarray(case.mesh.links, dtype=np.uint32).ravel()
        self.tags = np.ascontiguousarray(case.mesh.tags, dtype=np.uint8).ravel()
        self.affine = parallelogram(self.corners)
        self.bed = np.ascontiguousarray(case.bed, dtype=np.float64)
        self.manning = np.ascontiguousarray(case.manning, dtype=np.float64)
        # friction, wet/dry and subcells, and the incoming-wave tide where the case has it
        self.flags = (7 | (8 if getattr(case, "incoming_wave", False) else 0)) if flags is None else flags

    def _p(self, a):
        return a.ctypes.data_as(_dp)

    def stage(self, base, inp, a, b, c, time, rhs_only=False, fstar=None):
        c_ = self.case
        B = np.concatenate(base)
        I = np.concatenate(inp)
        out = np.empty_like(I)
        eta = c_.tide.elevation(time)
        lib().emu_stage(self.order, self.K, self._p(self.corners), self.links.ctypes.data, self.tags.ctypes.data,
                        self.affine.ctypes.data, self._p(self.bed), self._p(self.manning), self._p(B), self._p(I),
                        self._p(out), a, b, c, eta, c_.g, c_.h_min, c_.h_dry, 20.0, 1.0, c_.f0, int(rhs_only),
                        self.flags, None if fstar is None else self._p(fstar))
        return np.split(out, 3)

    def face_fluxes(self, h, hu, hv, t):
        out = np.empty(self.K * 4 * (self.order + 1) * 3)
        self.stage((h, hu, hv), (h, hu, hv), 0.0, 0.0, 0.0, t, rhs_only=True, fstar=out)
        return out.reshape(self.K, 4, self.order + 1, 3)

    def rhs(self, h, hu, hv, t):
        return self.stage((h, hu, hv), (h, hu, hv), 0.0, 0.0, 0.0, t, rhs_only=True)

    def element_dt(self, h, hu, hv, cfl, scheme=0):
        n = self.order
        ssp = 1.0 if scheme == 0 else 2.0
        lin = ([1.6, 1.04, 0.73, 0.56] if scheme == 0 else [2.2, 1.51, 1.15, 0.93])[n - 1]
        sub = ([0.94, 0.65, 0.47, 0.36] if scheme == 0 else [1.5, 1.1, 0.97, 0.75])[n - 1]
        pcfl = ssp * ((2.0 * n + 1.0) / (2.0 * n * (n + 1.0)))
        S = np.concatenate((h, hu, hv))
        out = np.empty(self.K)
        dt = lib().emu_dt(n, self.K, self._p(S), self._p(self.corners), self.links.ctypes.data,
                          self.affine.ctypes.data, cfl, self.case.g, self.case.h_min, self.case.h_dry, pcfl,
                          max(pcfl, 0.9 * lin), max(pcfl, 0.9 * sub), int(bool(self.flags & 2)), self._p(out))
        return out, dt

    def stage_where(self, base, inp, out, acc, elements, plan):
        """A pass of local time stepping (wet_lts_kernel and the subcell kernel's
        LTS instantiation) with each level's c and tide elevation from
        element_stage; out and acc (lists of 3 arrays) are updated in place."""
        from dgcom_amd import multirate as M
        c_ = self.case
        B = np.concatenate(base)
        I = np.concatenate(inp)
        O = np.concatenate(out)
        A = np.concatenate(acc) if acc is not None else None
        lc = np.zeros(M_LEVELS + 1)
        le = np.zeros(M_LEVELS + 1)
        for l in range(M_LEVELS + 1):
            lv = min(l, plan.finest)
            es = M.element_stage((plan.t, plan.dt), plan.finest, plan.substep, plan.depth, plan.weights, False, lv, lv)
            lc[l] = es.c
            le[l] = c_.tide.elevation(es.time)
        el = np.ascontiguousarray(elements, dtype=np.uint32)
        level = np.ascontiguousarray(plan.level, dtype=np.uint8)
        lrate = np.ascontiguousarray(plan.rate, dtype=np.uint8)
        w = plan.weights
        clips = lib().emu_stage_where(self.order, self.K, self._p(self.corners), self.links.ctypes.data,
                                      self.tags.ctypes.data, self.affine.ctypes.data, self._p(self.bed),
                                      self._p(self.manning), self._p(B), self._p(I), self._p(O),
                                      None if A is None else self._p(A), el.ctypes.data, el.size, level.ctypes.data,
                                      lrate.ctypes.data, plan.depth, w.start, w.input, self._p(lc), self._p(le), c_.g,
                                      c_.h_min, c_.h_dry, 20.0, 1.0, c_.f0, self.flags)
        for dst, src in zip(out, np.split(O, 3)):
            dst[...] = src
        if acc is not None:
            for dst, src in zip(acc, np.split(A, 3)):
                dst[...] = src
        return clips

    def coriolis_nodes(self, beta, y0):
        """f0 + beta (y - y0) at every node, by the library's coriolis_node_kernel."""
        f = np.empty(self.bed.size)
        lib().emu_coriolis_nodes(self.order, self.K, self._p(self.corners), self.case.f0, beta, y0, self._p(f))
        return f

    def stage_where_forced(self, base, inp, out, acc, elements, plan, bt, manning_node, coriolis_node):
        """stage_where through the kernels' forced instantiations: tides per boundary
        face node `bt` (a cases.BoundaryTides) with each level's moment row evaluated here
        as the library's host does (the ramp, and f(t)/f0 and the C library's sincos of
        omega' t per constituent, at the level's stage time), Manning n and Coriolis f per node."""
        from dgcom_amd import multirate as M, tides as T
        c_ = self.case
        B = np.concatenate(base)
        I = np.concatenate(inp)
        O = np.concatenate(out)
        A = np.concatenate(acc) if acc is not None else None
        ROW = lib().emu_lvl_moment()
        lc = np.zeros(M_LEVELS + 1)
        le = np.zeros(M_LEVELS + 1)
        rows = np.zeros((M_LEVELS + 1, ROW))
        for l in range(M_LEVELS + 1):
            lv = min(l, plan.finest)
            es = M.element_stage((plan.t, plan.dt), plan.finest, plan.substep, plan.depth, plan.weights, False, lv, lv)
            lc[l] = es.c
            le[l] = c_.tide.elevation(es.time)
            time = es.time
            if bt.ramp > 0.0:
                tau = min(max(time / bt.ramp, 0.0), 1.0)
                rows[l, 0] = tau * tau * (3.0 - 2.0 * tau)
            else:
                rows[l, 0] = 1.0
            for j in range(bt.omega.size):
                sn, cs = T.sincos(bt.omega[j] * time)
                rows[l, 1 + 3 * j] = 1.0 + bt.f_rate[j] * time
                rows[l, 2 + 3 * j] = cs
                rows[l, 3 + 3 * j] = sn
        el = np.ascontiguousarray(elements, dtype=np.uint32)
        level = np.ascontiguousarray(plan.level, dtype=np.uint8)
        lrate = np.ascontiguousarray(plan.rate, dtype=np.uint8)
        slot = np.ascontiguousarray(bt.slot, dtype=np.uint32).ravel()
        coef, mean, src = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (bt.coef, bt.mean,
                                                                                         bt.source_depth))
        mn = np.ascontiguousarray(manning_node, dtype=np.float64)
        cf = np.ascontiguousarray(coriolis_node, dtype=np.float64)
        w = plan.weights
        clips = lib().emu_stage_where_forced(
            self.order, self.K, self._p(self.corners), self.links.ctypes.data, self.tags.ctypes.data,
            self.affine.ctypes.data, self._p(self.bed), self._p(self.manning), self._p(B), self._p(I), self._p(O),
            None if A is None else self._p(A), el.ctypes.data, el.size, level.ctypes.data, lrate.ctypes.data,
            plan.depth, w.start, w.input, self._p(lc), self._p(le), c_.g, c_.h_min, c_.h_dry, 20.0, 1.0, c_.f0,
            self.flags, slot.ctypes.data, self._p(coef), self._p(mean), self._p(src), int(bt.omega.size),
            int(bt.velocity), bt.max_ratio, self._p(rows), self._p(mn), self._p(cf))
        for dst, src_ in zip(out, np.split(O, 3)):
            dst[...] = src_
        if acc is not None:
            for dst, src_ in zip(acc, np.split(A, 3)):
                dst[...] = src_
        return clips

    def finish_where(self, state, acc, elements):
        S = np.concatenate(state)
        A = np.concatenate(acc)
        el = np.ascontiguousarray(elements, dtype=np.uint32)
        clips = lib().emu_finish_where(self.order, self.K, self._p(self.corners), self.affine.ctypes.data, self._p(S),
                                       self._p(A), el.ctypes.data, el.size, self.case.h_dry, 20.0,
                                       int(bool(self.flags & 2)))
        for dst, src in zip(state, np.split(S, 3)):
            dst[...] = src
        return clips

    def stage_vjp(self, base, inp, w, a, b, c, time, rhs_only=False):
        """Cotangents of (base, input, bed, Manning n) of a stage for the output cotangent w."""
        c_ = self.case
        B = np.concatenate(base)
        I = np.concatenate(inp)
        W = np.concatenate(w)
        gb = np.zeros_like(B)
        gi = np.zeros_like(I)
        gbed = np.zeros(self.bed.size)
        gman = np.zeros(self.K)
        eta = c_.tide.elevation(time)
        lib().emu_stage_vjp(self.order, self.K, self._p(self.corners), self.links.ctypes.data, self.tags.ctypes.data,
                            self.affine.ctypes.data, self._p(self.bed), self._p(self.manning), self._p(B), self._p(I),
                            self._p(W), self._p(gb), self._p(gi), self._p(gbed), self._p(gman), a, b, c, eta, c_.g,
                            c_.h_min, c_.h_dry, 20.0, 1.0, c_.f0, int(rhs_only), self.flags)
        return np.split(gb, 3), np.split(gi, 3), gbed, gman


def multirate_assign(links, n1, max_levels, n_stages, line_rates, edts, flags, wet_dry, dt_caps, landings,
                     device=False):
    """The library's level assignment (multirate::Assigner) over a sequence of
    calls: per call (dt, stable, finest, level[K], rates[n_stages][K]); with
    `device`, through the device half the library runs (emulated kernels)."""
    links = np.ascontiguousarray(links, dtype=np.uint32).ravel()
    K = links.size // 4
    calls = len(edts)
    E = np.ascontiguousarray(np.concatenate(edts), dtype=np.float64)
    F = np.ascontiguousarray(np.concatenate(flags), dtype=np.uint8)
    cap = np.array([c if c is not None else 0.0 for c in dt_caps], dtype=np.float64)
    land = np.array([x if x is not None else 0.0 for x in landings], dtype=np.float64)
    dt, stable = np.zeros(calls), np.zeros(calls)
    finest = np.zeros(calls, dtype=np.int32)
    level = np.zeros(calls * K, dtype=np.uint8)
    rates = np.zeros(calls * n_stages * K, dtype=np.uint8)
    P = lambda a: a.ctypes.data_as(_dp)
    fn = lib().emu_multirate_assign_device if device else lib().emu_multirate_assign
    fn(K, n1, links.ctypes.data, max_levels, n_stages, int(line_rates), calls, P(E),
                               F.ctypes.data, int(wet_dry), P(cap), P(land), P(dt), P(stable), finest.ctypes.data,
                               level.ctypes.data, rates.ctypes.data)
    return [(dt[q], stable[q], int(finest[q]), level[q * K:(q + 1) * K],
             rates[q * n_stages * K:(q + 1) * n_stages * K].reshape(n_stages, K)) for q in range(calls)]
