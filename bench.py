"""Benchmark of the 2D shallow-water `WetDry` SSP-RK3 time step on one MI355X.

Workload (the large unstructured configuration): 1,000,000 quads from a
1000 x 1000 grid over 100 km x 100 km with interior vertices jittered by up to
20 % of the cell size (seeded), order N = 3 (16 nodes per element), bed =
-(40 m + a seeded Gaussian random field of sd 25 m and 5 km correlation), a
1 m M2 tide on the west boundary, Manning n = 0.025, f = 1.3e-4 s^-1, SSP-RK3
at CFL 0.5 with the positivity-relaxed time step, from still water.

A step is one adaptive SSP-RK3 time step (the time-step kernel, its global
minimum read back to the host, three fused stage kernels), inputs resident in
HBM. value = DOF-updates/s = 3 * K * (N+1)^2 * 3 stages * steps / wall time.
Before the warm-up the library's autotuner (dg_tune) picks the stage's launch
shape for this order and size, untimed (--no-tune keeps the default).

Besides the timed loop: the stage kernel's average launch duration from HIP
events on the library's stream gives the roofline point (algorithmic bytes per
element-stage, SoA FP64: 8*NN*(3 input + 3 base if base != input + 1 bed +
3 out) + 64 corners + 16 face links + 8 Manning n); rank 0 also times the CPU
restatement of the reference (OpenMP, all host threads) on a bounded sample of
the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from dgcom_amd import cases, swe2d  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
STAGES = 3


def stage_bytes(nn: int, base_differs: bool) -> int:
    """Algorithmic HBM bytes of one element-stage."""
    return 8 * nn * (3 + (3 if base_differs else 0) + 1 + 3) + 64 + 16 + 8


def read_traffic(workload: str):
    """HBM bytes per stage (wet + subcell kernels) measured with rocprofv3's
    FETCH_SIZE and WRITE_SIZE counters, scaled by what they read for a plain copy
    of known size, if that summary is in profiles/traffic_<workload>.json."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_stage_launch")
    return None


def read_fp64(workload: str):
    """FP64 FLOPs per stage launch (wet + subcell kernels) from rocprofv3's FP64 VALU
    instruction counters, if that summary is in profiles/fp64_<workload>.json."""
    path = os.path.join(ROOT, "profiles", f"fp64_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)["per_dispatch"]
    flops = sum(v["fp64_flops"] for k, v in d.items() if "stage_kernel" in k)
    return flops, d


def cpu_baseline(case, seconds_target: float):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as checker

    threads = checker.max_threads()
    orc = checker.Oracle.from_case(case)
    h, hu, hv = case.h.copy(), case.hu.copy(), case.hv.copy()
    # one warm step, then as many as fit the target (at least one)
    t, sec = orc.run_timed(h, hu, hv, 0.0, 1, case.cfl)
    n = max(1, min(50, int(seconds_target / max(sec, 1e-9))))
    t, sec = orc.run_timed(h, hu, hv, t, n, case.cfl)
    dofs = 3 * case.n_elements * case.nn * STAGES * n
    try:
        model = [line.split(":", 1)[1].strip() for line in open("/proc/cpuinfo") if line.startswith("model name")][0]
    except Exception:
        model = "unknown"
    return {"value": dofs / sec, "unit": "DOF-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} SSP-RK3 steps of the full workload after 1 warm step, {sec:.1f} s, OpenMP "
                      f"{threads} threads on {model}"}


def adjoint_bench(n, n_steps, snapshots, every, n_stations):
    """The adjoint workload: wall time of the station-misfit gradient (forward
    sweep + reverse sweep with recomputation from `snapshots` checkpoints) over
    the wall time of one plain forward run of the same steps, the reverse
    sweep alone over the forward when every state is kept (per-step VJP cost),
    the gradient of the same schedule with step records (snapshots + 1 records of
    stage values and RHS: no reverse step recomputes its stages), and the
    reverse sweep with every step recorded over the first 400 steps against their
    plain forward."""
    from dgcom_amd import inversion

    t0 = time.time()
    inv = inversion.workload(n=n, n_steps=n_steps, every=every, n_stations=n_stations)
    c = inv.case
    phys = swe2d.SWEPhysics2D.from_case(c)
    inversion.observe(inv, phys)
    setup = time.time() - t0
    x0 = phys.state(c.h, c.hu, c.hv)

    def timed(fn):
        phys.sync()
        w0 = time.perf_counter()
        r = fn()
        phys.sync()
        return r, time.perf_counter() - w0

    plain = phys.tape(1)
    J, fwd_s = timed(lambda: plain.forward(x0, 0.0, inv.dts, every, inv.obs))
    J, fwd_s = timed(lambda: plain.forward(x0, 0.0, inv.dts, every, inv.obs))
    plain.close()
    tape = phys.tape(snapshots)
    Jg, f1 = timed(lambda: tape.forward(x0, 0.0, inv.dts, every, inv.obs))
    (gbed, gman), b1 = timed(lambda: tape.backward(1.0))
    fsteps, vjps = tape.stats()
    tape.close()
    # the same schedule with step records (snapshots + 1 of them): no reverse step recomputes its stages
    tape = phys.tape(snapshots)
    tape.keep_stages()
    Jr, f4 = timed(lambda: tape.forward(x0, 0.0, inv.dts, every, inv.obs))
    _, b4 = timed(lambda: tape.backward(1.0))
    used = tape.records_used()
    tape.close()
    per_step = None
    if n_steps <= 4000:
        full = phys.tape(n_steps + 1)
        _, f2 = timed(lambda: full.forward(x0, 0.0, inv.dts, every, inv.obs))
        _, b2 = timed(lambda: full.backward(1.0))
        full.close()
        per_step = {"forward_with_all_checkpoints_s": f2, "reverse_sweep_s": b2, "ratio": b2 / fwd_s,
                    "gradient_over_forward_keeping_every_state": (f2 + b2) / fwd_s}
    # step records (every step's stage values and RHS kept, nothing recomputed) over
    # the first steps that fit HBM at 6 states per step
    n_rec = min(n_steps, 400)
    dts_r, obs_r = inv.dts[:n_rec], inv.obs[: n_rec // every]
    short = phys.tape(1)
    _, fr = timed(lambda: short.forward(x0, 0.0, dts_r, every, obs_r))
    _, fr = timed(lambda: short.forward(x0, 0.0, dts_r, every, obs_r))
    short.close()
    rec = phys.tape(n_rec + 1)
    rec.keep_stages()
    _, f3 = timed(lambda: rec.forward(x0, 0.0, dts_r, every, obs_r))
    _, b3 = timed(lambda: rec.backward(1.0))
    rec.close()
    records = {"steps": n_rec, "forward_s": fr, "forward_keeping_stages_s": f3, "reverse_sweep_s": b3,
               "ratio": b3 / fr, "gradient_over_forward": (f3 + b3) / fr}
    K, nn = c.n_elements, c.nn
    return {"workload": f"bathymetry inversion: {K} jittered quads ({n}x{n} over 50 km), N={c.order}, GRF truth bed, "
                        f"prior = truth + GRF(sd 2 m), {n_stations} stations every {every} steps, {n_steps} SSP-RK3 "
                        f"steps, {snapshots} checkpoints",
            "elements": K, "order": c.order, "steps": n_steps, "snapshots": snapshots,
            "forward_s": fwd_s, "gradient_s": f1 + b1, "ratio": (f1 + b1) / fwd_s,
            "with_step_records": {"gradient_s": f4 + b4, "ratio": (f4 + b4) / fwd_s, "forward_sweep_s": f4,
                                  "reverse_sweep_s": b4, "reverse_steps_from_records": used,
                                  "misfit_equal": Jr == Jg},
            "gradient_forward_sweep_s": f1, "gradient_reverse_sweep_s": b1,
            "forward_steps": fsteps, "step_vjps": vjps, "misfit": Jg,
            "grad_norms": {"bed": float(np.linalg.norm(gbed)), "manning": float(np.linalg.norm(gman))},
            "per_step": per_step, "step_records": records, "setup_s": setup}


def adjoint_ensemble_bench(n_members, n_steps, snapshots):
    """The inversion ensemble (inversion.ensemble_workload: uneven meshes, N in
    {2, 3}, each member with its own truth/prior beds and 12 stations): every
    member's misfit gradient from its own tape of `snapshots` states, the members'
    sweeps in one sequence of launches, against the members' plain forward runs of
    the same steps one after another."""
    from dgcom_amd import inversion

    t0 = time.time()
    members = inversion.ensemble_workload(n_members=n_members, n_steps=n_steps)
    physics = []
    for inv in members:
        phys = swe2d.SWEPhysics2D.from_case(inv.case)
        inversion.observe(inv, phys)
        physics.append(phys)
    setup = time.time() - t0
    fwd = 0.0
    for inv, phys in zip(members, physics):
        c = inv.case
        x0 = phys.state(c.h, c.hu, c.hv)
        plain = phys.tape(1)
        phys.sync()
        w0 = time.perf_counter()
        plain.forward(x0, 0.0, inv.dts, inv.every, inv.obs)
        phys.sync()
        fwd += time.perf_counter() - w0
        plain.close()
        x0.free()
    physics[0].sync()
    w0 = time.perf_counter()
    grads = inversion.ensemble_gradients(members, physics, snapshots)
    for phys in physics:
        phys.sync()
    grad = time.perf_counter() - w0
    w0 = time.perf_counter()
    grads_r = inversion.ensemble_gradients(members, physics, snapshots, keep_stages=True)
    for phys in physics:
        phys.sync()
    grad_r = time.perf_counter() - w0
    same = all(a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
               for a, b in zip(grads, grads_r))
    orders = [inv.case.order for inv in members]
    return {"workload": f"inversion ensemble: {n_members} members, K log-uniform 20000-250000 jittered quads, N in "
                        f"{{2,3}}, own GRF truth/prior beds (land in each), 12 stations every 10 steps, {n_steps} "
                        f"SSP-RK3 steps, {snapshots} checkpoints per member",
            "members": n_members, "elements_total": int(sum(inv.case.n_elements for inv in members)),
            "members_per_order": {str(o): orders.count(o) for o in (2, 3)},
            "forward_runs_s": fwd, "gradients_s": grad, "ratio": grad / fwd,
            "with_step_records": {"gradients_s": grad_r, "ratio": grad_r / fwd, "same_gradients": bool(same)},
            "misfits": [float(J) for J, _, _ in grads], "setup_s": setup}


def ensemble_bench(n_members, n_steps, k_lo, k_hi):
    """An ensemble of uneven members advanced together (one launch per order per
    kernel, graph-replayed stages, tuned launch shapes) against the same members
    run one after another (one run per launch sequence)."""
    t0 = time.time()
    members = cases.ensemble_members(n_members, 99, k_lo, k_hi, n_steps)
    phys = [swe2d.SWEPhysics2D.from_case(c) for c in members]
    setup = time.time() - t0
    dofs = sum(3 * c.n_elements * c.nn * STAGES * n_steps for c in members)
    # one run after another
    seq = 0.0
    for c, p in zip(members, phys):
        s = p.state(c.h, c.hu, c.hv)
        p.sync()
        w0 = time.perf_counter()
        p.run(s, 0.0, n_steps, c.cfl)
        p.sync()
        seq += time.perf_counter() - w0
        s.free()
    ens = swe2d.Ensemble(phys)
    for m, c in enumerate(members):
        ens.upload(m, c.h, c.hu, c.hv)
    ens.run(1, members[0].cfl, graph=True, tune=True)  # tuning and graph capture, untimed
    for m, c in enumerate(members):
        ens.upload(m, c.h, c.hu, c.hv)
    ens.t[:] = 0.0
    ens.members[0].sync()
    w0 = time.perf_counter()
    ens.run(n_steps, members[0].cfl, graph=True)
    ens.members[0].sync()
    batched = time.perf_counter() - w0
    orders = [c.order for c in members]
    return {"workload": f"{n_members} members, K log-uniform {k_lo}-{k_hi} jittered quads, N uniform in {{1,2,3}}, "
                        f"M2 tide 0.2-1.5 m with random phase, GRF beds, {n_steps} SSP-RK3 steps each",
            "members": n_members, "elements_total": int(sum(c.n_elements for c in members)),
            "members_per_order": {str(o): orders.count(o) for o in (1, 2, 3)},
            "dof_updates": dofs, "batched_s": batched, "one_run_per_launch_s": seq,
            "batched_dof_updates_per_s": dofs / batched, "one_run_per_launch_dof_updates_per_s": dofs / seq,
            "speedup": seq / batched, "tuning": {str(o): ens.tuning(o) for o in (1, 2, 3) if o in orders},
            "setup_s": setup}


def c1_bench():
    """The idealised tidal estuary (100 x 20 affine quads, N = 2, 0.5 m M2 tide, 25 h from rest): the steps that
    reach 25 h, timed on the GPU and in the CPU restatement on all host threads, and whether the two end states
    are equal bit for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as checker

    c = cases.tidal_estuary()
    phys = swe2d.SWEPhysics2D.from_case(c)
    s = phys.state(c.h, c.hu, c.hv)
    _, dts = phys.run(s, 0.0, 100_000, c.cfl, record_dt=True)
    n = int(np.searchsorted(np.cumsum(dts), 25 * 3600.0)) + 1
    s.upload(c.h, c.hu, c.hv)
    phys.sync()
    w0 = time.perf_counter()
    t_gpu = phys.run(s, 0.0, n, c.cfl)
    phys.sync()
    gpu_s = time.perf_counter() - w0
    got = s.download()
    orc = checker.Oracle.from_case(c)
    h, hu, hv = c.h.copy(), c.hu.copy(), c.hv.copy()
    t_cpu, cpu_s = orc.run_timed(h, hu, hv, 0.0, n, c.cfl)
    equal = t_gpu == t_cpu and all(np.array_equal(a, b) for a, b in zip(got, (h, hu, hv)))
    dofs = 3 * c.n_elements * c.nn * STAGES * n
    phys.close()
    return {"workload": "idealised tidal estuary: 100x20 affine quads over 10 km x 2 km, N=2, depth 10 -> 2 m, "
                        "0.5 m M2 tide (1 h ramp), Manning 0.025, f=1.2e-4, SSP-RK3 CFL 0.5, 25 h from rest",
            "steps": n, "model_time_s": t_gpu, "gpu_s": gpu_s, "gpu_ms_per_step": gpu_s / n * 1e3,
            "gpu_dof_updates_per_s": dofs / gpu_s, "cpu_s": cpu_s, "cpu_dof_updates_per_s": dofs / cpu_s,
            "cpu_cores": checker.max_threads(), "cpu_kind": "port", "bitwise_equal": bool(equal)}


def lts_bench(nx, ny, t_end, max_levels):
    """Local time stepping against global steps over the same model time: the
    reference's tidal channel refined to nx x ny elements (N = 2, SSP-RK(4,3) at
    its default CFL, wet/dry flats, M2 + S2 with nodal corrections), from still
    water for t_end seconds at up to `max_levels` levels. Global steps: the
    runner's loop (one dt and one fused step per step, landing on t_end). Local:
    the library's multirate stepper (levels and stage rates from the device's
    element time steps each coarse step, then its passes). Both from the same
    state, warm."""
    c = cases.tidal_channel(nx=nx, ny=ny)
    phys = swe2d.SWEPhysics2D.from_case(c)
    s = phys.state(c.h, c.hu, c.hv)
    sim = swe2d.Simulation(phys, c.cfl, c.scheme)
    sim.run(s, 0.0, min(60.0, t_end))                      # warm-up (library, caches)
    s.upload(c.h, c.hu, c.hv)
    phys.sync()
    w0 = time.perf_counter()
    _, n_glob = sim.run(s, 0.0, t_end)
    phys.sync()
    glob = time.perf_counter() - w0
    h_glob = s.download()[0]
    s.upload(c.h, c.hu, c.hv)
    warm = swe2d.Multirate(phys, max_levels, c.scheme)
    warm.run(s, 0.0, min(60.0, t_end), c.cfl)                        # warm-up
    warm.close()
    mr = swe2d.Multirate(phys, max_levels, c.scheme)
    s.upload(c.h, c.hu, c.hv)
    phys.sync()
    w0 = time.perf_counter()
    _, n_lts = mr.run(s, 0.0, t_end, c.cfl)
    phys.sync()
    lts = time.perf_counter() - w0
    h_lts = s.download()[0]
    stats = mr.stats()
    mr.close()
    s.free()
    m = cases.node_masses(c)
    glob_evals = 4 * c.n_elements * n_glob
    phys.close()
    return {"workload": f"tidal channel refined to {c.n_elements} elements ({nx}x{ny}, N=2), {t_end:.0f} s from "
                        f"still water, SSP-RK(4,3) CFL {c.cfl:.3f}, up to {max_levels} levels",
            "global_s": glob, "global_steps": n_glob, "lts_s": lts, "coarse_steps": n_lts,
            "finest_level": stats["finest_level"], "speedup": glob / lts,
            "global_element_evaluations": glob_evals, "lts_element_evaluations": stats["element_evaluations"],
            "evaluations_ratio": glob_evals / stats["element_evaluations"],
            "lts_element_evaluations_per_s": stats["element_evaluations"] / lts,
            "global_element_evaluations_per_s": glob_evals / glob,
            "volume_difference_rel": abs(float(m @ h_lts) - float(m @ h_glob)) / float(m @ h_glob)}


def weather_bench(case, phys, state, t, n_steps, ms_plain):
    """The main workload's step under gridded weather (GriddedAtmosphere2D's source: wind stress and pressure
    gradient from two snapshots an hour apart, linear in time, at every node): its cost over the plain step."""
    x, y = cases.node_coordinates(case.mesh, case.order)
    x, y = x.ravel(), y.ravel()

    def fields(xx, yy, tt):
        u = 8.0 + 6.0 * np.sin(xx / 7000.0 + tt / 3000.0)
        return u, 4.0 * np.cos(yy / 9000.0), 1e-4 * np.sin(xx / 20000.0), 1e-4 * np.cos(yy / 15000.0), 0.0 * xx

    t0 = time.time()
    wx = cases.weather_at_nodes(case, [t - 1.0, t + 3600.0], fields)
    phys.set_weather(wx)
    setup = time.time() - t0
    t = phys.run(state, t, 2, case.cfl)
    phys.sync()
    w0 = time.perf_counter()
    t = phys.run(state, t, n_steps, case.cfl)
    phys.sync()
    ms = (time.perf_counter() - w0) / n_steps * 1e3
    phys.set_weather(None)
    return {"workload": f"the main workload under gridded weather: {x.size} nodes x 2 snapshots x 5 fields f32 "
                        f"({wx.series.nbytes / 1e6:.0f} MB resident), Large & Pond drag, linear in time",
            "steps": n_steps, "ms_per_step": ms, "ms_per_step_plain": ms_plain, "ratio": ms / ms_plain,
            "setup_s": setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1000, help="grid side (1000: the 1M-element workload)")
    ap.add_argument("--fma", action="store_true", help="the FMA-contracted variant of the library")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape (no dg_tune)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--adjoint-n", type=int, default=500, help="grid side of the adjoint workload (500: 250,000 quads)")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--adjoint-snapshots", type=int, default=20)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-c1", action="store_true", help="skip the 25 h tidal estuary (GPU and CPU)")
    ap.add_argument("--adjoint-ensemble-members", type=int, default=32,
                    help="members of the inversion ensemble's gradient timing (0: not run)")
    ap.add_argument("--adjoint-ensemble-steps", type=int, default=1000)
    ap.add_argument("--lts", default="800x80,2000x200", help="refined tidal channels for the local-time-stepping "
                    "leg, comma-separated nx x ny ('none' for none)")
    ap.add_argument("--lts-seconds", type=float, default=900.0)
    ap.add_argument("--lts-levels", type=int, default=8)
    ap.add_argument("--weather-steps", type=int, default=0,
                    help="steps of the main workload under gridded weather (wind, pressure, 2 snapshots; 0: not run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the time step runs on a single device here "
              "(element partitions with a face-trace halo per stage are the multi-GPU design); "
              "--gpus must be 1", file=sys.stderr)
        sys.exit(2)

    t0 = time.time()
    case = cases.large_unstructured(n=args.n)
    gen_s = time.time() - t0
    K, nn = case.n_elements, case.nn
    phys = swe2d.SWEPhysics2D.from_case(case, fma=args.fma)
    state = phys.state(case.h, case.hu, case.hv)
    # the stage's launch shape for this order and size (dg_tune, untimed; no result changes)
    shape = phys.tune(state) if not args.no_tune else phys.launch_shape()

    t = phys.run(state, 0.0, args.warmup, case.cfl) if args.warmup > 0 else 0.0
    phys.sync()
    w0 = time.perf_counter()
    t = phys.run(state, t, args.steps, case.cfl)
    phys.sync()
    wall = time.perf_counter() - w0
    assert state.is_finite(), "the state is not finite"
    dofs = 3 * K * nn * STAGES * args.steps
    value = dofs / wall

    # Roofline point of the stage kernel: HIP events around each launch (separate pass)
    phys.timing(True)
    t = phys.run(state, t, args.profile_steps, case.cfl)
    stage_ms, launches, dt_ms, dt_launches = phys.timing_read()
    phys.timing(False)
    steps_p = launches // STAGES
    bytes_total = steps_p * K * (stage_bytes(nn, False) + 2 * stage_bytes(nn, True))
    achieved = bytes_total / (stage_ms * 1e-3) / 1e9
    copy_bytes = 2 * 3 * K * nn * 8
    copy_ms = phys.copy_bench(3 * K * nn * 8, reps=20)
    copy_gbs = copy_bytes / (copy_ms * 1e-3) / 1e9
    nd, es = phys.masks(state)
    workload = f"large_unstructured_n{args.n}_N3"
    traffic = read_traffic(workload)
    fp64 = read_fp64(workload)

    out = {
        "metric": "SWE-2D RK-stage DOF-updates/s + %HBM roofline; adjoint/forward time ratio",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered mesh and Gaussian-random-field bed)",
        "config": {"workload": f"large unstructured: {K} jittered quads ({args.n}x{args.n} over 100 km), N=3, "
                               f"GRF bed, 1 m M2 tide, Manning 0.025, f=1.3e-4, SSP-RK3 CFL 0.5",
                   "elements": K, "order": 3, "nodes": K * nn, "scheme": "ssp_rk3", "stages_per_step": STAGES,
                   "library": "fma" if args.fma else "exact",
                   "subcell_elements_at_end": int(es.sum()), "dry_nodes_at_end": int(nd.sum()),
                   "model_time_s": t},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "one SSP-RK3 stage: wet_stage_kernel<3> + list compaction + subcell_stage_kernel<3>",
                     "avg_launch_ms": stage_ms / max(launches, 1),
                     "bytes_per_launch_avg": bytes_total / max(launches, 1),
                     "measured_copy_GBps": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "dt_kernel_avg_ms": dt_ms / max(dt_launches, 1),
                     "launch_shape": {**shape, "tuned": not args.no_tune}},
        "setup_s": gen_s,
    }
    if fp64 is not None:
        flops, per = fp64
        tf = flops / (stage_ms / max(launches, 1) * 1e-3) / 1e12
        out["fp64_valu"] = {"achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS,
                            "flops_per_stage_launch": flops,
                            "wet_kernel_valu_insts_per_wave": per.get("wet_stage_kernel<3>", {}).get("valu_insts_per_wave"),
                            "wet_kernel_fp64_insts_per_wave": per.get("wet_stage_kernel<3>", {}).get("fp64_insts_per_wave")}
    progress("the large unstructured workload", t0)
    if args.weather_steps > 0:
        out["weather"] = weather_bench(case, phys, state, t, args.weather_steps, out["ms_per_step"])
        progress("the weather leg", t0)
    if not args.no_adjoint:
        phys.close()
        out["adjoint"] = adjoint_bench(args.adjoint_n, args.adjoint_steps, args.adjoint_snapshots, 10, 20)
        progress("the adjoint", t0)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(args.ensemble_members, args.ensemble_steps, 20_000, 400_000)
        progress("the ensemble", t0)
    if args.adjoint_ensemble_members > 0:
        out["adjoint_ensemble"] = adjoint_ensemble_bench(args.adjoint_ensemble_members, args.adjoint_ensemble_steps, 20)
        progress("the inversion ensemble", t0)
    if args.lts and args.lts != "none":
        out["lts"] = []
        for spec in args.lts.split(","):
            out["lts"].append(lts_bench(*map(int, spec.split("x")), args.lts_seconds, args.lts_levels))
            progress(f"local time stepping on the {spec} channel", t0)
    if not args.no_c1:
        out["c1"] = c1_bench()
        progress("the 25 h tidal estuary", t0)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
        progress("the CPU baseline", t0)
    print(json.dumps(out), flush=True)


def progress(what: str, t0: float):
    """One line on stderr per finished leg (the JSON result is the only stdout line)."""
    print(f"bench: {what} done ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
