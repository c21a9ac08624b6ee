"""Discrete adjoint of the batched RK45 solve (device entry bg_rk_vjp_device).

`solve_vjp` returns, for every instance, the loss
    L = sum_k |y(t_k) - target_k|^2
over shared observation times t_k, its gradient with respect to the
per-instance parameters and initial state, and the states at the observation
times.  `RK45Loss` wraps it in a torch.autograd.Function (used for gradient
checks): its forward runs the forward and backward GPU passes, its backward
scales the stored gradients by the incoming cotangent.  A `Schedule` carries
the longest-first order from one solve to the next in a loop.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L


def _ptr(t):
    return None if t is None else t.data_ptr()


class Schedule:
    """The forward pass's schedule carried from one solve of an ensemble to
    the next, as in an optimisation loop: after each solve it holds the
    longest-first order by that solve's attempt counts (`longest_first`),
    and the next solve of the same n instances runs in that order.  Results
    do not depend on the order."""

    def __init__(self):
        self.order = None

    def current(self, n: int):
        return self.order if self.order is not None and self.order.numel() == n else None

    def update(self, counts: torch.Tensor, stream=None):
        self.order = longest_first(counts, stream)


def _vjp_call(rhs, y0, params, obs_t, targets, t_end, tol, dt_min, dt_max, t0,
              checkpoint_every, max_checkpoints, stream, order, blocks_per_cu, tuning=None,
              plan=None):
    """Builds the adjoint descriptor and output tensors, then runs
    bg_rk_vjp_device, bg_adjoint_autotune_device when `tuning` (an
    L.AdjointTuning to fill) is given, or bg_adjoint_solve_tuned_device with
    `plan` (a filled one)."""
    dev = y0.device
    D, n = y0.shape
    P = params.shape[0]
    f64 = dict(dtype=torch.float64, device=dev)
    y0, params = y0.contiguous(), params.contiguous()
    obs_t, targets = obs_t.contiguous(), targets.contiguous()
    te = t_end if torch.is_tensor(t_end) else None
    tl = tol if torch.is_tensor(tol) else None
    out = dict(
        loss=torch.zeros(n, **f64), grad_params=torch.zeros((P, n), **f64),
        grad_y0=torch.zeros((D, n), **f64), obs_y=torch.zeros((obs_t.numel(), D, n), **f64),
        t=torch.zeros(n, **f64), y=torch.zeros((D, n), **f64),
        status=torch.zeros(n, dtype=torch.int32, device=dev),
        counts=torch.zeros((n, 2), dtype=torch.int32, device=dev))
    base = L.EnsembleDesc(
        n=n, dim=D, n_params=P, y0=_ptr(y0), params=_ptr(params),
        t0_s=t0, t_end=_ptr(te), t_end_s=0.0 if te is not None else float(t_end),
        tol=_ptr(tl), tol_s=0.0 if tl is not None else float(tol),
        dt_min_s=dt_min, dt_max_s=dt_max, t_out=_ptr(out["t"]), y_out=_ptr(out["y"]),
        status=_ptr(out["status"]), counts=_ptr(out["counts"]),
        order=_ptr(order) if order is not None else None)
    d = L.AdjointDesc(base=base, n_obs=obs_t.numel(), checkpoint_every=checkpoint_every,
                      max_checkpoints=max_checkpoints, blocks_per_cu=blocks_per_cu,
                      obs_t=_ptr(obs_t),
                      targets=_ptr(targets), loss=_ptr(out["loss"]),
                      grad_params=_ptr(out["grad_params"]), grad_y0=_ptr(out["grad_y0"]),
                      obs_y=_ptr(out["obs_y"]))
    sh = C.c_void_p(stream.cuda_stream if stream is not None
                    else torch.cuda.current_stream(dev).cuda_stream)
    if plan is not None:
        rc = L.lib().bg_adjoint_solve_tuned_device(L.RK45, L.RHS[rhs], C.byref(d),
                                                   C.byref(plan), sh)
        name = "bg_adjoint_solve_tuned_device"
    elif tuning is None:
        rc = L.lib().bg_rk_vjp_device(L.RK45, L.RHS[rhs], C.byref(d), sh)
        name = "bg_rk_vjp_device"
    else:
        rc = L.lib().bg_adjoint_autotune_device(L.RK45, L.RHS[rhs], C.byref(d),
                                                C.byref(tuning), sh)
        name = "bg_adjoint_autotune_device"
    if rc != 0:
        raise RuntimeError(f"{name} returned {rc}")
    return out


def solve_vjp(rhs: str, y0: torch.Tensor, params: torch.Tensor, obs_t: torch.Tensor,
              targets: torch.Tensor, t_end, tol, dt_min: float = 1e-12, dt_max: float = 0.1,
              t0: float = 0.0, checkpoint_every: int = 16, max_checkpoints: int = 128,
              stream=None, order=None, blocks_per_cu: int = 0, schedule: Schedule = None,
              plan: "L.AdjointTuning" = None):
    """All tensors float64 on the GPU: y0 [D, n], params [P, n], obs_t [n_obs],
    targets [n_obs, D, n]; t_end / tol scalars or [n] tensors.  `order` (int32
    [n], optional) is the forward pass's schedule, e.g. from
    bg_rk_schedule_device with a plan tuned on the same problem.  With a
    `Schedule` instead, the solve uses the order it carries (if any) and
    leaves it the order for the next solve.  `plan` (from `autotune_vjp`)
    replaces checkpoint_every, max_checkpoints and blocks_per_cu and, without
    an `order`, may order the forward pass by its fitted cost model
    (bg_adjoint_solve_tuned_device)."""
    if schedule is not None and order is None:
        order = schedule.current(y0.shape[1])
    out = _vjp_call(rhs, y0, params, obs_t, targets, t_end, tol, dt_min, dt_max, t0,
                    checkpoint_every, max_checkpoints, stream, order, blocks_per_cu, plan=plan)
    if schedule is not None:
        schedule.update(out["counts"], stream)
    return out


def autotune_vjp(rhs: str, y0: torch.Tensor, params: torch.Tensor, obs_t: torch.Tensor,
                 targets: torch.Tensor, t_end, tol, dt_min: float = 1e-12, dt_max: float = 0.1,
                 t0: float = 0.0, checkpoint_every: int = 16, max_checkpoints: int = 128,
                 stream=None, order=None, blocks_per_cu: int = 0) -> "L.AdjointTuning":
    """The adjoint's plan (bg_adjoint_autotune_device): the fastest checkpoint
    interval, resident blocks per CU and (without an `order`) forward order,
    input or fitted longest-first, for this problem, timed on the whole
    gradient solve; the arguments are solve_vjp's.  Returns L.AdjointTuning
    (pass it to solve_vjp as `plan`)."""
    tn = L.AdjointTuning()
    _vjp_call(rhs, y0, params, obs_t, targets, t_end, tol, dt_min, dt_max, t0,
              checkpoint_every, max_checkpoints, stream, order, blocks_per_cu, tuning=tn)
    return tn


def longest_first(counts: torch.Tensor, stream=None) -> torch.Tensor:
    """Schedule for a later solve of the same instances: the permutation
    (int32 [n], device) that runs them longest first by the attempts of an
    earlier solve (`counts` [n, 2] int32: accepted, rejected), computed by
    bg_schedule_by_counts_device.  In an optimisation loop the previous
    iteration's counts predict the next one's; pass the result as `order`."""
    counts = counts.contiguous()
    n = counts.shape[0]
    order = torch.empty(n, dtype=torch.int32, device=counts.device)
    sh = C.c_void_p(stream.cuda_stream if stream is not None
                    else torch.cuda.current_stream(counts.device).cuda_stream)
    rc = L.lib().bg_schedule_by_counts_device(_ptr(counts), n, _ptr(order), sh)
    if rc != 0:
        raise RuntimeError(f"bg_schedule_by_counts_device returned {rc}")
    return order


class RK45Loss(torch.autograd.Function):
    """loss[i] = sum_k |y_i(t_k) - target_ik|^2 with gradients w.r.t. params and y0."""

    @staticmethod
    def forward(ctx, params, y0, rhs, obs_t, targets, t_end, tol, schedule=None):
        r = solve_vjp(rhs, y0.detach(), params.detach(), obs_t, targets, t_end, tol,
                      schedule=schedule)
        ctx.save_for_backward(r["grad_params"], r["grad_y0"])
        return r["loss"]

    @staticmethod
    def backward(ctx, g):
        gp, gy = ctx.saved_tensors
        return gp * g[None, :], gy * g[None, :], None, None, None, None, None, None
