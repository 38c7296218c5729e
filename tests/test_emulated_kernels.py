"""The product's HIP kernel source, run on the CPU through an emulation of
workgroups (tests/emu: one host thread per lane, a barrier per __syncthreads),
against the CPU restatement of the reference: RHS, face fluxes, stages, element
time steps and whole SSP-RK3 steps, bit for bit (both built without
multiply-add contraction). This pins the kernels' arithmetic and indexing on a
machine without a GPU; the GPU tests then pin the device build.
"""
import numpy as np
import pytest

import emu_bridge as E
import oracle as O
from cases_for_tests import CASES, perturbed


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_rhs_and_face_fluxes_bitwise(name, make):
    case = make()
    orc = O.Oracle.from_case(case)
    ek = E.EmulatedKernels(case)
    for t in (0.0, 1800.0):
        h, hu, hv = perturbed(case, seed=int(t))
        for g, r in zip(ek.rhs(h, hu, hv, t), orc.rhs(h, hu, hv, t)):
            assert np.array_equal(g, r)
        fg = ek.face_fluxes(h, hu, hv, t)[..., 0]
        fo = orc.face_mass(h, hu, hv, t)
        assert np.array_equal(fg, fo)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_element_dt_bitwise(name, make):
    case = make()
    orc = O.Oracle.from_case(case)
    ek = E.EmulatedKernels(case)
    h, hu, hv = perturbed(case, seed=3)
    for scheme in (0, 1):
        got, dmin = ek.element_dt(h, hu, hv, case.cfl, scheme)
        ref = orc.element_dt(h, hu, hv, case.cfl, scheme)
        assert np.array_equal(got, ref)
        assert dmin == orc.dt(h, hu, hv, case.cfl, scheme)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_ssp_rk3_steps_bitwise(name, make):
    case = make()
    orc = O.Oracle.from_case(case)
    ek = E.EmulatedKernels(case)
    h, hu, hv = case.h.copy(), case.hu.copy(), case.hv.copy()
    s = [h.copy(), hu.copy(), hv.copy()]
    t = 0.0
    for _ in range(4):
        dt = orc.dt(h, hu, hv, case.cfl)
        assert ek.element_dt(*s, case.cfl)[1] == dt
        inp = s
        for a, b, beta, c in ((1.0, 0.0, 1.0, 0.0), (0.75, 0.25, 0.25, 1.0), (1 / 3, 2 / 3, 2 / 3, 0.5)):
            inp = ek.stage(s, inp, a, b, beta * dt, t + c * dt)
        s = inp
        orc.step(h, hu, hv, t, dt)
        t += dt
        for g, r in zip(s, (h, hu, hv)):
            assert np.array_equal(g, r)
