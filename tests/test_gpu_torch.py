"""The torch.autograd.Function wrappers: gradients through torch's autograd
equal the C ABI's VJPs (stage and step exactly; a chain of steps against the
checkpointed misfit gradient to rounding), and a gradient step on the bed
lowers the station misfit."""
import numpy as np
import pytest
import torch

from dgcom_amd import autograd, cases, stations, swe2d
import oracle as O
from cases_for_tests import perturbed, small_unstructured

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")


def test_stage_autograd_equals_the_vjp():
    c = cases.drying_beach(order=2, nx=8, ny=3)
    base = perturbed(c, seed=1)
    dt = O.Oracle.from_case(c).dt(*base, c.cfl)
    phys = swe2d.SWEPhysics2D.from_case(c)
    xb = _t(np.stack(base)).requires_grad_(True)
    xi = _t(np.stack(base)).requires_grad_(True)
    bed = _t(c.bed).requires_grad_(True)
    man = _t(c.manning).requires_grad_(True)
    y = autograd.stage(phys, xb, xi, bed, man, 0.75, 0.25, 0.25 * dt, 0.5 * dt)
    # the forward equals the plain stage
    ref = phys.state()
    phys.stage_where(phys.state(*base), phys.state(*base), ref, 0.75, 0.25, 0.25 * dt, 0.5 * dt)
    assert np.array_equal(y.detach().cpu().numpy(), np.stack(ref.download()))
    rng = np.random.default_rng(2)
    w = rng.standard_normal((3, c.h.size))
    (y * _t(w)).sum().backward()
    gb, gi = phys.state().zero(), phys.state().zero()
    phys.param_grads_zero()
    phys.stage_vjp(phys.state(*base), phys.state(*base), 0.75, 0.25, 0.25 * dt, 0.5 * dt, phys.state(*w), gb, gi)
    gbed, gman = phys.param_grads()
    assert np.array_equal(xb.grad.cpu().numpy(), np.stack(gb.download()))
    assert np.array_equal(xi.grad.cpu().numpy(), np.stack(gi.download()))
    assert np.array_equal(bed.grad.cpu().numpy(), gbed)
    assert np.array_equal(man.grad.cpu().numpy(), gman)


def _problem():
    c = small_unstructured(2)
    st = stations.random_stations(c, 5, seed=3, min_depth=2.0)
    x0 = perturbed(c, seed=4, amp=0.02)
    phys = swe2d.SWEPhysics2D.from_case(c)
    _, dts = phys.run(phys.state(*x0), 0.0, 20, c.cfl, record_dt=True)
    x, y = cases.node_coordinates(c.mesh, c.order)
    truth = c.bed - 0.5 * cases.gaussian_random_field(x, y, (0.0, 2400.0, 0.0, 2400.0), 1.0, 400.0, 8).ravel()
    phys.set_bed(truth)
    phys.set_stations(st.elements, st.weights)
    s = phys.state(*x0)
    obs, t = [], 0.0
    for i, dt in enumerate(dts):
        phys.step(s, t, dt)
        t += dt
        if (i + 1) % 4 == 0:
            obs.append(phys.sample_stations(s))
    phys.set_bed(c.bed)
    return c, st, x0, dts, np.array(obs), phys


def test_step_chain_matches_the_checkpointed_misfit_gradient():
    c, st, x0, dts, obs, phys = _problem()
    bed = _t(c.bed).requires_grad_(True)
    man = _t(c.manning).requires_grad_(True)
    x = _t(np.stack(x0)).requires_grad_(True)
    el = torch.tensor(st.elements[:, None] * phys.nn + np.arange(phys.nn)[None, :], device="cuda")
    w = _t(st.weights)
    J, s, t = 0.0, x, 0.0
    for i, dt in enumerate(dts):
        s = autograd.step(phys, s, bed, man, t, dt)
        t += dt
        if (i + 1) % 4 == 0:
            eta = (w * (s[0][el] + bed[el])).sum(1)
            J = J + ((eta - _t(obs[(i + 1) // 4 - 1])) ** 2).sum()
    J.backward()
    J_tape = autograd.station_misfit(phys, _t(c.bed).requires_grad_(True), _t(c.manning), _t(np.stack(x0)), 0.0, dts,
                                     4, obs, 3)
    assert abs(J_tape.item() - J.item()) <= 1e-12 * J.item()
    bed2 = _t(c.bed).requires_grad_(True)
    man2 = _t(c.manning).requires_grad_(True)
    x2 = _t(np.stack(x0)).requires_grad_(True)
    autograd.station_misfit(phys, bed2, man2, x2, 0.0, dts, 4, obs, 3).backward()
    for a, b in ((bed.grad, bed2.grad), (man.grad, man2.grad), (x.grad, x2.grad)):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    # the tape with step records: the same gradients bit for bit
    bed3 = _t(c.bed).requires_grad_(True)
    man3 = _t(c.manning).requires_grad_(True)
    x3 = _t(np.stack(x0)).requires_grad_(True)
    autograd.station_misfit(phys, bed3, man3, x3, 0.0, dts, 4, obs, 3, keep_stages=True).backward()
    for a, b in ((bed3.grad, bed2.grad), (man3.grad, man2.grad), (x3.grad, x2.grad)):
        assert torch.equal(a, b)


def test_gradient_step_lowers_the_misfit():
    c, st, x0, dts, obs, phys = _problem()
    bed = _t(c.bed).requires_grad_(True)
    man = _t(c.manning)
    x = _t(np.stack(x0))
    J0 = autograd.station_misfit(phys, bed, man, x, 0.0, dts, 4, obs, 4)
    J0.backward()
    g = bed.grad
    lr = 1e-3 * J0.item() / float((g * g).sum())
    J1 = autograd.station_misfit(phys, (bed - lr * g).detach().contiguous(), man, x, 0.0, dts, 4, obs, 4)
    assert J1.item() < J0.item()
