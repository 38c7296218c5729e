"""torch.autograd.Function wrappers over the C ABI's forward / VJP pairs.

Tensors are float64 on the GPU: a state is (3, K*NN) (h, hu, hv rows, SoA),
the bed (K*NN,), Manning n (K,). The library runs on torch's current stream
and reads and writes the tensors' memory in place (dg_state_wrap), so nothing
is copied through the host.

- `stage(phys, base, input, bed, manning, a, b, c, time)`: one fused SSP stage
  (dg_stage); backward = dg_stage_vjp. Saved for backward: the stage's inputs.
- `step(phys, x, bed, manning, t, dt, scheme)`: one SSP step (dg_step);
  backward = dg_step_vjp, which recomputes the stage values from the saved x.
- `station_misfit(phys, bed, manning, x0, t0, dts, every, obs, snapshots)`:
  J = sum (eta - obs)^2 over the station samples of a run at the given dts;
  forward = the first sweep of a binomial checkpoint schedule (the tape of
  `snapshots` states is what is saved for backward), backward = the reverse
  sweep (dg_misfit_backward).
The reference has no autograd; the state layout and the stage are those of
`SWESolution2D` and `stage_where` (src/physics/builder.rs:652-737).
"""
from __future__ import annotations

import numpy as np
import torch

from . import swe2d


def _prepare(phys, bed, manning):
    for t in (bed, manning):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("bed and manning must be contiguous float64 CUDA tensors")
    phys.set_stream(torch.cuda.current_stream(bed.device).cuda_stream)
    phys.set_bed(device_ptr=bed.data_ptr())
    phys.set_manning(device_ptr=manning.data_ptr())


def _state_tensor(phys, x):
    if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous() or x.shape != (3, phys.n_nodes_total):
        raise ValueError(f"a state must be a contiguous float64 CUDA tensor of shape (3, {phys.n_nodes_total})")
    return phys.wrap(x.data_ptr())


class _Stage(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phys, base, inp, bed, manning, a, b, c, time):
        _prepare(phys, bed, manning)
        out = torch.empty_like(inp)
        phys.stage_where(_state_tensor(phys, base), _state_tensor(phys, inp), _state_tensor(phys, out), a, b, c, time)
        ctx.save_for_backward(base, inp, bed, manning)
        ctx.phys = phys
        ctx.coef = (a, b, c, time)
        return out

    @staticmethod
    def backward(ctx, gout):
        base, inp, bed, manning = ctx.saved_tensors
        phys = ctx.phys
        _prepare(phys, bed, manning)
        gout = gout.contiguous()
        gbase, ginp = torch.zeros_like(base), torch.zeros_like(inp)
        gbed, gman = torch.empty_like(bed), torch.empty_like(manning)
        phys.param_grads_zero()
        a, b, c, time = ctx.coef
        phys.stage_vjp(_state_tensor(phys, base), _state_tensor(phys, inp), a, b, c, time, _state_tensor(phys, gout),
                       _state_tensor(phys, gbase), _state_tensor(phys, ginp))
        phys.param_grads_to_device(gbed.data_ptr(), gman.data_ptr())
        return None, gbase, ginp, gbed, gman, None, None, None, None


class _Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phys, x, bed, manning, t, dt, scheme):
        _prepare(phys, bed, manning)
        out = x.detach().clone()
        phys.step(_state_tensor(phys, out), t, dt, scheme)
        ctx.save_for_backward(x, bed, manning)
        ctx.phys = phys
        ctx.args = (t, dt, scheme)
        return out

    @staticmethod
    def backward(ctx, gout):
        x, bed, manning = ctx.saved_tensors
        phys = ctx.phys
        _prepare(phys, bed, manning)
        gout = gout.contiguous()
        gx = torch.empty_like(x)
        gbed, gman = torch.empty_like(bed), torch.empty_like(manning)
        phys.param_grads_zero()
        t, dt, scheme = ctx.args
        phys.step_vjp(_state_tensor(phys, x), t, dt, _state_tensor(phys, gout), _state_tensor(phys, gx), scheme)
        phys.param_grads_to_device(gbed.data_ptr(), gman.data_ptr())
        return None, gx, gbed, gman, None, None, None


class _StationMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phys, bed, manning, x0, t0, dts, every, obs, snapshots, scheme, keep_stages):
        _prepare(phys, bed, manning)
        tape = phys.tape(snapshots)
        if keep_stages:
            tape.keep_stages()
        J = tape.forward(_state_tensor(phys, x0), t0, dts, every, obs, scheme)
        ctx.tape = tape
        ctx.phys = phys
        ctx.save_for_backward(bed, manning, x0)
        return torch.tensor(J, dtype=torch.float64, device=bed.device)

    @staticmethod
    def backward(ctx, gJ):
        bed, manning, x0 = ctx.saved_tensors
        phys = ctx.phys
        _prepare(phys, bed, manning)
        gx0 = torch.empty_like(x0)
        gbed, gman = torch.empty_like(bed), torch.empty_like(manning)
        ctx.tape.backward(float(gJ), _state_tensor(phys, gx0))
        phys.param_grads_to_device(gbed.data_ptr(), gman.data_ptr())
        ctx.tape.close()
        return None, gbed, gman, gx0, None, None, None, None, None, None, None


def stage(phys, base, inp, bed, manning, a, b, c, time):
    return _Stage.apply(phys, base, inp, bed, manning, a, b, c, time)


def step(phys, x, bed, manning, t, dt, scheme=swe2d.SSP_RK3):
    return _Step.apply(phys, x, bed, manning, t, dt, scheme)


def station_misfit(phys, bed, manning, x0, t0, dts, every, obs, snapshots, scheme=swe2d.SSP_RK3, keep_stages=False):
    """Requires `phys.set_stations` first; obs is [len(dts) // every][stations].
    keep_stages: the tape keeps step records (no reverse step recomputes its
    stages; 5 states per SSP-RK3 record, snapshots + 1 of them)."""
    return _StationMisfit.apply(phys, bed, manning, x0, t0, np.asarray(dts, dtype=np.float64), every,
                                np.asarray(obs, dtype=np.float64), snapshots, scheme, bool(keep_stages))
