"""Multirate local time stepping: the host side that drives a physics module's
element-subset stages (src/time/multirate.rs).

Every element gets a level l and takes 2^l substeps of dt/2^l per coarse step
dt (Constantinescu & Sandu's multirate partitioned Runge-Kutta on an SSP base
scheme). The stepper assigns the levels from the element time steps, limits
neighbouring levels to differ by one, spreads the stage rates (the finest level
an element's stage-s RHS reads, one face stencil per stage), and runs the
coarse step as passes: per finest substep, each stage over the elements whose
rate reaches the substep's depth, then the end of the substeps ending there.

The physics module ("local") provides, as `LocalTimeStepping` does
(multirate.rs:296-391): `element_dt_state(state, cfl, scheme)`,
`depth_flags(state, d0, d1, d2)` (for `coarse_step_bound`),
`begin_coarse_step(level, rates)`, `stage_where(base, input, elements, plan,
out, acc)`, `finish_where(state, acc, elements, depth)`, and the state
operations `clone_state` / `copy_state`. `swe2d.SWEPhysics2D` runs them on the
GPU through the C ABI (`dg_begin_coarse_step`, `dg_stage_where`,
`dg_finish_where`); the same stepper drives any other implementation of them.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

MAX_LEVELS = 16                     # multirate.rs:455
LANDING_SLACK = 1e-9                # multirate.rs:588
SUBCELL_CAP_DEPTH_FACTOR = 100.0    # swe_2d.rs:2185
LINE_LOCAL_HEADROOM = 1.0           # builder.rs:148

SSP_RK3, SSP_RK43 = 0, 1


@dataclasses.dataclass(frozen=True)
class ShuOsherStage:
    """Stage s + 1 = start*S + input*Y_s + beta*dt*F(Y_s), at time t + c*dt
    (integrator.rs:209-263)."""
    start: float
    input: float
    beta: float
    c: float


STAGES = {
    SSP_RK3: (ShuOsherStage(1.0, 0.0, 1.0, 0.0), ShuOsherStage(0.75, 0.25, 0.25, 1.0),
              ShuOsherStage(1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 0.5)),
    SSP_RK43: (ShuOsherStage(1.0, 0.0, 0.5, 0.0), ShuOsherStage(0.0, 1.0, 0.5, 0.5),
               ShuOsherStage(2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 1.0), ShuOsherStage(0.0, 1.0, 0.5, 0.5)),
}


@dataclasses.dataclass(frozen=True)
class ElementStage:
    time: float
    a: float
    b: float
    c: float
    accumulate: tuple | None   # (w, keep) for the last stage


def element_stage(step, finest, substep, depth, stage: ShuOsherStage, last, level, rate) -> ElementStage:
    """The stage of an element on `level` whose stage RHS follows `rate`, in
    finest substep `substep` of depth `depth` (multirate.rs:169-200)."""
    t, dt = step
    dt_l = dt / float(1 << level)
    q = substep >> (finest - level)
    t_k = t + float(q) * dt_l
    acc = None
    if last:
        acc = (1.0 / float(1 << (rate - level)), 0.0 if level >= depth else 1.0)
    return ElementStage(t_k + stage.c * dt_l, stage.start, stage.input, stage.beta * dt_l, acc)


@dataclasses.dataclass
class StagePlan:
    """One stage pass of a coarse step (multirate.rs:203-243)."""
    t: float
    dt: float
    finest: int
    substep: int
    depth: int
    stage: int
    weights: ShuOsherStage
    last: bool
    level: np.ndarray
    rate: np.ndarray

    def at(self, k: int) -> ElementStage:
        return element_stage((self.t, self.dt), self.finest, self.substep, self.depth, self.weights, self.last,
                             int(self.level[k]), int(self.rate[k]))


def neighbours(links: np.ndarray) -> np.ndarray:
    """Face neighbours [K][4] of a face-link table, -1 on boundary faces."""
    links = np.asarray(links, dtype=np.uint32)
    nb = (links >> 4).astype(np.int64)
    return np.where(links == 0xFFFFFFFF, -1, nb)


def assign_levels(element_dt, max_levels: int, dt_cap=None):
    """The coarse step 2^L min(element_dt), L the most levels up to
    `max_levels` keeping it at most max(element_dt) and `dt_cap`, and each
    element's level, the smallest l with dt/2^l <= its step (multirate.rs:526-568).
    Returns (dt, L, levels); (inf, 0, zeros) when no element limits the step."""
    assert max_levels <= MAX_LEVELS
    edt = np.asarray(element_dt, dtype=np.float64)
    levels = np.zeros(edt.size, dtype=np.uint8)
    finite = edt[np.isfinite(edt)]
    if finite.size == 0:
        return math.inf, 0, levels
    dt_min, dt_max = float(finite.min()), float(finite.max())
    finest = 0
    while finest < max_levels:
        coarser = dt_min * float(2 << finest)
        if coarser > dt_max or (dt_cap is not None and coarser > dt_cap):
            break
        finest += 1
    dt = dt_min * float(1 << finest)
    dt_l = np.full(edt.size, dt)
    for l in range(finest):
        refine = dt_l > edt
        if not refine.any():
            break
        levels[refine] += 1
        dt_l[refine] *= 0.5
    return dt, finest, levels


def levels_for_step(element_dt, dt: float, max_levels: int):
    """The smallest l <= max_levels with dt/2^l <= each element's step, and the
    finest used (multirate.rs:574-586)."""
    edt = np.asarray(element_dt, dtype=np.float64)
    levels = np.zeros(edt.size, dtype=np.uint8)
    dt_l = np.full(edt.size, dt)
    for _ in range(max_levels):
        refine = dt_l > edt
        if not refine.any():
            break
        levels[refine] += 1
        dt_l[refine] *= 0.5
    return levels, int(levels.max()) if levels.size else 0


def spread_to_neighbours(nb: np.ndarray, frm: np.ndarray, drop: int):
    """to[k] = max(from[k], from[j] - drop) over k's face neighbours j, and
    whether any changed (multirate.rs:592-613)."""
    frm = np.asarray(frm, dtype=np.int16)
    to = frm.copy()
    for f in range(4):
        j = nb[:, f]
        has = j >= 0
        cand = np.where(has, np.maximum(frm[np.where(has, j, 0)] - drop, 0), 0)
        to = np.maximum(to, cand)
    to = to.astype(np.uint8)
    return to, bool((to != frm).any())


def face_nodes(n1: int) -> np.ndarray:
    """Node indices [4][n1] of the four faces, counter-clockwise (node j*n1 + i
    on r-line j and s-line i; DGOperators2D::face_nodes): node fi of a face
    coincides with node n1 - 1 - fi of the neighbour's."""
    fi = np.arange(n1)
    return np.stack([fi, fi * n1 + n1 - 1, (n1 - 1) * n1 + (n1 - 1 - fi), (n1 - 1 - fi) * n1])


def inner_offset(face: int, n1: int) -> int:
    """The node one in from a face node, normal to the face (multirate.rs:274-283)."""
    return (n1, -1, -n1, 1)[face]


def spread_over_lines(nb: np.ndarray, links: np.ndarray, n1: int, line_local: np.ndarray, frm: np.ndarray):
    """One stage's node-by-node spreading (multirate.rs:705-767): `frm` the node
    rates [K][n1^2] of the stage's input; returns the node rates of its value and
    each element's RHS rate. A line-local element's node reads its two GLL lines
    and, on a face, the neighbour's coincident node; any other element reads all
    its own nodes and each neighbour's face nodes and the nodes one in from them."""
    K, nn = frm.shape
    fn = face_nodes(n1)
    links = np.asarray(links, dtype=np.uint32)
    nbf = (links & 3).astype(np.int64)
    # whole elements
    finest = frm.max(axis=1)
    for f in range(4):
        has = nb[:, f] >= 0
        j = np.where(has, nb[:, f], 0)
        g = nbf[:, f]
        # the neighbour's nodes of its face g and the nodes one in from them
        nodes = fn[g]                                            # [K][n1]
        inner = nodes + np.array([inner_offset(q, n1) for q in range(4)])[g][:, None]
        vals = np.maximum(frm[j[:, None], nodes], frm[j[:, None], inner]).max(axis=1)
        finest = np.where(has, np.maximum(finest, vals), finest)
    to = np.repeat(finest[:, None], nn, axis=1).astype(np.uint8)
    # line-local elements
    own = frm.reshape(K, n1, n1)                                 # [k][j][i]
    line = np.maximum(own.max(axis=2)[:, :, None], own.max(axis=1)[:, None, :]).reshape(K, nn)
    for f in range(4):
        has = nb[:, f] >= 0
        j = np.where(has, nb[:, f], 0)
        g = nbf[:, f]
        other = fn[g][:, ::-1]                                   # the neighbour's node coinciding with fi
        vals = frm[j[:, None], other]
        cur = line[:, fn[f]]
        line[:, fn[f]] = np.where(has[:, None], np.maximum(cur, vals), cur)
    ll = np.asarray(line_local, dtype=bool)
    to[ll] = line[ll]
    rate = to.max(axis=1).astype(np.uint8)
    return to, rate


def limit_level_jumps(nb: np.ndarray, level: np.ndarray) -> np.ndarray:
    """Raise levels until face neighbours differ by at most one (multirate.rs:815-821)."""
    level = np.asarray(level, dtype=np.uint8)
    while True:
        to, changed = spread_to_neighbours(nb, level, 1)
        if not changed:
            return level
        level = to


def at_or_above(key: np.ndarray, max_key: int):
    """lists[d] = the elements with key >= d, in increasing order (multirate.rs:825-838)."""
    key = np.asarray(key)
    return [np.nonzero(key >= d)[0].astype(np.uint32) for d in range(max_key + 1)]


def depth(p: int, finest: int) -> int:
    """The coarsest level whose substeps start at finest substep p
    (multirate.rs:842-849)."""
    if p == 0 or p == 1 << finest:
        return 0
    return finest - ((p & -p).bit_length() - 1)


@dataclasses.dataclass
class MultirateStats:
    """Work counts of a multirate run (multirate.rs:856-889)."""
    steps: int = 0
    element_evaluations: int = 0
    global_evaluations: int = 0
    finest_level: int = 0
    model_seconds: float = 0.0

    def speedup(self) -> float:
        return self.global_evaluations / max(self.element_evaluations, 1)


class MultirateStepper:
    """Levels, stage rates, pass lists and stage storage of a multirate run
    (MultirateStepper, multirate.rs:713-1330; the face stencil)."""

    def __init__(self, max_levels: int, scheme: int = SSP_RK3, line_rates: bool = True):
        assert max_levels <= MAX_LEVELS
        self.line_rates = line_rates
        self.memo = None
        self.scheme = scheme
        self.stages = STAGES[scheme]
        self.max_levels = max_levels
        self.finest = 0
        self.level = None
        self.rate = []
        self.active = []
        self.ending = []
        self.stable_step = math.inf
        self.evaluations_per_step = 0
        self.values = None
        self.last = None
        self.acc = None
        self.stats = MultirateStats()

    # -- levels ------------------------------------------------------------
    def assign_levels(self, local, nb: np.ndarray, state, cfl: float, dt_cap=None, landing=None) -> float:
        """Levels from the current element time steps (multirate.rs:1014-1043),
        the coarse step at most the module's bound and `dt_cap`, fitted to
        `landing` (the time left to a time the run must reach)."""
        edt = local.element_dt_state(state, cfl, self.scheme)
        flags = self.depth_flags(local, state)
        bound = self.coarse_step_bound(local, edt, flags)
        cap = bound if dt_cap is None else min(dt_cap, bound)
        lines = None
        if self.line_rates and flags is not None and getattr(local, "n1", 0) >= 2:
            lines = (local.n1, local.links, (flags & 4) != 0)
        return self.assign(edt, nb, cap if math.isfinite(cap) else None, landing, lines)

    @staticmethod
    def depth_flags(local, state):
        """Bit 0: every node deeper than SUBCELL_CAP_DEPTH_FACTOR h_dry; bit 1:
        than h_dry (the coarse-step bound); bit 2: than the line-local depth,
        h_dry + LINE_LOCAL_HEADROOM (SWEPhysics2D::line_local, builder.rs:563-594,
        the depth where the WetDry kernel changes)."""
        h_dry = getattr(local, "line_local_h_dry", None)
        if h_dry is None:
            return None
        return local.depth_flags(state, SUBCELL_CAP_DEPTH_FACTOR * h_dry, h_dry, h_dry + LINE_LOCAL_HEADROOM)

    @staticmethod
    def coarse_step_bound(local, element_dt, flags) -> float:
        """SWEPhysics2D::coarse_step_bound (builder.rs:614-637): the largest step
        of an element whose every node is deeper than SUBCELL_CAP_DEPTH_FACTOR
        h_dry, else than h_dry; none without a positivity bound."""
        if getattr(local, "positivity_h_dry", None) is None or flags is None:
            return math.inf
        edt = np.asarray(element_dt)
        fin = np.isfinite(edt)
        for bit in (1, 2):
            sel = fin & ((flags & bit) != 0)
            longest = float(edt[sel].max()) if sel.any() else 0.0
            if longest > 0.0:
                return longest
        return math.inf

    def assign(self, element_dt, nb: np.ndarray, dt_cap=None, landing=None, lines=None) -> float:
        """multirate.rs:1072-1168 with the face stencil; with `lines` = (n1,
        face links, line-local flags) the rates are spread node by node."""
        dt, finest, level = assign_levels(element_dt, self.max_levels, dt_cap)
        self.stable_step = dt
        if landing is not None and finest > 0 and landing > 0.0 and math.isfinite(dt):
            steps = max(math.ceil(landing / (dt * (1.0 + LANDING_SLACK))), 1.0)
            fitted = landing / steps
            if fitted < dt:
                dt = fitted
                level, finest = levels_for_step(element_dt, dt, self.max_levels)
        self.finest = finest
        # the limit, the rates and the lists stand while their inputs do
        # (multirate.rs:1114-1121)
        key = (finest, level.tobytes(), None if lines is None else np.asarray(lines[2], dtype=bool).tobytes())
        if self.memo is not None and self.memo == key:
            return dt
        level = limit_level_jumps(nb, level)
        rates = []
        if lines is not None:
            n1, links, ll = lines
            frm = np.repeat(level[:, None], n1 * n1, axis=1)
            for _ in self.stages:
                frm, rate = spread_over_lines(nb, links, n1, ll, frm)
                rates.append(rate)
        else:
            frm = level
            for _ in self.stages:
                frm, _ = spread_to_neighbours(nb, frm, 0)
                rates.append(frm)
        self.memo = key
        self.level, self.rate = level, rates
        self.evaluations_per_step = int(sum((np.left_shift(1, r.astype(np.int64))).sum() for r in rates))
        self.active = [at_or_above(r, finest) for r in rates]
        self.ending = at_or_above(level, finest)
        return dt

    def assign_one_level(self, n_elements: int):
        """Every element on one level: global steps through the element stages
        (multirate.rs:1205-1218)."""
        self.memo = None
        self.finest = 0
        self.level = np.zeros(n_elements, dtype=np.uint8)
        self.rate = [self.level.copy() for _ in self.stages]
        self.evaluations_per_step = len(self.stages) * n_elements
        self.active = [at_or_above(r, 0) for r in self.rate]
        self.ending = at_or_above(self.level, 0)

    def planned_speedup(self) -> float:
        glob = len(self.stages) * self.level.size * (1 << self.finest)
        return glob / max(self.evaluations_per_step, 1)

    # -- the coarse step ---------------------------------------------------
    def schedule(self, t: float, dt: float, pass_fn):
        """The passes of one coarse step, in order (multirate.rs:1260-1296)."""
        finest = self.finest
        n_stages = len(self.stages)
        for p in range(1 << finest):
            d = depth(p, finest)
            for s, w in enumerate(self.stages):
                pass_fn("stage", s, self.active[s][d],
                        StagePlan(t, dt, finest, p, d, s, w, s + 1 == n_stages, self.level, self.rate[s]))
            nxt = depth(p + 1, finest)
            pass_fn("finish", nxt, self.ending[nxt], None)

    def step(self, local, state, t: float, dt: float):
        """One coarse step of dt from t with the levels of the last assignment
        (multirate.rs:1298-1330)."""
        n_stages = len(self.stages)
        if self.values is None:
            self.values = [local.clone_state(state) for _ in range(n_stages - 1)]
            self.last = local.clone_state(state)
            self.acc = local.clone_state(state)
        values, last, acc = self.values, self.last, self.acc
        local.begin_coarse_step(self.level, self.rate)

        def run(kind, s, elements, plan):
            if kind == "stage":
                inp = state if s == 0 else values[s - 1]
                if s + 1 < n_stages:
                    local.stage_where(state, inp, elements, plan, values[s], None)
                else:
                    local.stage_where(state, inp, elements, plan, last, acc)
            else:
                local.finish_where(state, acc, elements, s)

        self.schedule(t, dt, run)
        st = self.stats
        st.steps += 1
        st.element_evaluations += self.evaluations_per_step
        st.global_evaluations += n_stages * self.level.size * (1 << self.finest)
        st.finest_level = max(st.finest_level, self.finest)
        st.model_seconds += dt

    def release(self, local):
        for x in (self.values or []) + ([self.last, self.acc] if self.last is not None else []):
            free = getattr(local, "free_state", None)
            if free is not None:
                free(x)
        self.values = self.last = self.acc = None


def run(local, nb, state, t: float, t_end: float, cfl: float, max_levels: int = 8, scheme: int = SSP_RK3,
        dt_max=None, stepper: MultirateStepper | None = None, max_steps=None, interval=None, callback=None):
    """Simulation::run_with_callback with a multirate integrator
    (runner.rs:298-455): levels and the coarse step fitted to land on the next
    target (t_end, or with `interval` the next output time t_start + k interval,
    where `callback(state, t)` is called), then one coarse step, until t_end.
    Returns (t, coarse steps, stepper)."""
    stepper = stepper or MultirateStepper(max_levels, scheme)
    n, k, t_start = 0, 0, t
    while t < t_end:
        if max_steps is not None and n >= max_steps:
            raise RuntimeError(f"Maximum step limit ({max_steps}) reached")
        target = t_end if interval is None else min(t_start + (k + 1) * interval, t_end)
        fitted = stepper.assign_levels(local, nb, state, cfl, dt_max, target - t)
        dt = stepper.stable_step
        if dt_max is not None:
            dt = min(dt, dt_max)
        dt = min(dt, fitted)
        landed = t + dt * (1.0 + LANDING_SLACK) >= target
        if landed:
            dt = target - t
        stepper.step(local, state, t, dt)
        t = target if landed else t + dt
        n += 1
        if interval is not None and t >= t_start + (k + 1) * interval:
            k += 1
            if callback is not None:
                callback(state, t)
    return t, n, stepper
