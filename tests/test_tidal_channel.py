"""The reference's example case cases/tidal_channel/case.toml and its known
answer (notes/gpu.md:146): 25 h of M2 and S2 with the ramp and the nodal
corrections of a run from 2025-06-15 through the characteristic boundary at
rest, Coriolis at 63.5 N, Manning 0.025, drying flats, SSP-RK(4,3) at the
default CFL, outputs every 30 min (each step that reaches one lands on it):
5,765 steps and no clip; with `velocity = "incoming_wave"` (the tide as a wave
entering from still water) 5,783 steps.

The step count follows the whole run: every adaptive step's positivity-relaxed
bound over the flooding and draining flats, the tide's amplitude and phase at
every stage (the astronomy of `dgcom_amd.tides`), the landing on the 50 output
times. The CPU restatement is run here; the GPU run equals it bit for bit."""
import math

import numpy as np
import pytest

from dgcom_amd import cases
import oracle as O

T_END, INTERVAL, LANDING_SLACK = 25 * 3600.0, 1800.0, 1e-9   # multirate.rs:588


def run_oracle(c):
    orc = O.Oracle.from_case(c)
    x = [c.h.copy(), c.hu.copy(), c.hv.copy()]
    t, n, k = 0.0, 0, 0
    while t < T_END:
        target = min((k + 1) * INTERVAL, T_END)
        dt = orc.dt(*x, c.cfl, c.scheme)
        landed = t + dt * (1.0 + LANDING_SLACK) >= target
        if landed:
            dt = target - t
        orc.run_dts(*x, t, [dt], c.scheme)
        t = target if landed else t + dt
        n += 1
        if t >= (k + 1) * INTERVAL:
            k += 1
    return x, n, k, orc.clips()


@pytest.fixture(scope="module")
def case():
    return cases.tidal_channel()


def test_the_case_as_the_reference_resolves_it(case):
    assert case.n_elements == 160 and case.order == 2 and case.scheme == 1
    assert abs(case.cfl - 1.359) < 1e-12                        # 0.9 x 1.51, "1.36 at order 2"
    assert abs(case.f0 - 1.3051928564507e-4) < 1e-16            # 2 Omega sin 63.5
    assert abs(case.tide.omega[1] * 12 * 3600.0 - 2 * math.pi) < 1e-12   # S2: 12 h
    f_m2 = case.tide.nodal[0][0]
    assert 0.96 < f_m2 < 0.97                                   # the node near N = 0 in 2025
    assert case.bed.max() == 2.0 and case.bed.min() == -30.0


KNOWN_STEPS = {"at_rest": 5765, "incoming_wave": 5783}


@pytest.mark.parametrize("velocity", ["at_rest", "incoming_wave"])
def test_known_answer_steps(velocity):
    c = cases.tidal_channel(velocity=velocity)
    x, n, k, clips = run_oracle(c)
    assert k == 50
    assert n == KNOWN_STEPS[velocity]
    assert clips == 0
    assert np.isfinite(x[0]).all() and x[0].min() >= 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("velocity", ["at_rest", "incoming_wave"])
def test_gpu_run_equals_the_restatement(velocity):
    from dgcom_amd import swe2d
    case = cases.tidal_channel(velocity=velocity)
    x, n, _, _ = run_oracle(case)
    phys = swe2d.SWEPhysics2D.from_case(case)
    s = phys.state(case.h, case.hu, case.hv)
    t, steps = swe2d.Simulation(phys, case.cfl, scheme=case.scheme).run(s, 0.0, T_END, interval=INTERVAL)
    assert steps == n == KNOWN_STEPS[velocity] and t == T_END
    got = s.download()
    for a, b in zip(got, x):
        assert np.array_equal(a, b)
    assert phys.negative_depth_clips() == 0
