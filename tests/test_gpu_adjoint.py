"""GPU discrete adjoint (bg_rk_vjp_device) against independent references.

  * forward pass with observation stops: bit-identical to the oracle's
    DEVICE-profile solve with the same stops (states at every observation),
    and within 10x tol of the oracle's BACON profile (the reference's own
    arithmetic: no fused multiply-adds, nalgebra's norm, powf) at every
    observation of 256 instances;
  * gradients vs central finite differences (h = 1e-6) of the loss evaluated
    by replaying the frozen accepted-step sequence in the oracle: relative
    error below 1e-6 per instance on 64 instances drawn over the whole
    parameter range (the step sequence comes from the oracle's own forward
    solve, which the first test shows equals the GPU's);
  * gradients vs the oracle's forward-sensitivity solve of the augmented
    14-dimensional Lotka-Volterra system at tol 1e-12: the continuous and
    discrete sensitivities differ by the discretisation error, so the bar is
    1e-4 relative;
  * the torch.autograd.Function plumbing and the checkpoint-capacity error.
"""
import numpy as np
import pytest

import oracle as O
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-8
N = 64


def _t(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a),
                           dtype=dtype or torch.float64, device="cuda:0")


@pytest.fixture(scope="module")
def lv():
    """64 Lotka-Volterra instances of the adjoint workload (parameters
    log-uniform over [0.5, 2], y0 uniform over [0.5, 2]^2), targets = the GPU
    solution at the perturbed parameters."""
    import torch
    from bacon_amd.adjoint import solve_vjp
    w = W.lotka_volterra(N)
    lo, hi = np.log(0.5), np.log(2.0)
    lp = np.log(w["params"])
    # the sample spans the range: every parameter reaches both outer fifths
    assert np.all(lp.min(1) < lo + 0.2 * (hi - lo)) and np.all(lp.max(1) > hi - 0.2 * (hi - lo))
    obs = w["obs"]
    zeros = torch.zeros((obs.size, 2, N), dtype=torch.float64, device="cuda:0")
    r = solve_vjp("lotka_volterra", _t(w["y0"]), _t(w["target_params"]), _t(obs), zeros,
                  15.0, TOL)
    assert torch.all(r["status"] == 0)
    w["targets"] = r["obs_y"].cpu().numpy()
    return w


def _gpu(w, **kw):
    from bacon_amd.adjoint import solve_vjp
    r = solve_vjp("lotka_volterra", _t(w["y0"]), _t(w["params"]), _t(w["obs"]),
                  _t(w["targets"]), 15.0, TOL, **kw)
    return {k: v.cpu().numpy() for k, v in r.items()}


def _oracle_path(w, i, params=None):
    p = w["params"][:, i] if params is None else params
    return O.solve(O.RK45, "lotka_volterra", w["y0"][:, i], 0.0, 15.0, TOL, 1e-12, 0.1,
                   params=p, semantics=O.INTENDED, profile=O.PROFILE_DEVICE,
                   stops=w["obs"][:-1], path_cap=100000)


def test_forward_with_stops_bit_exact(gpu, orc, lv):
    g = _gpu(lv)
    assert np.all(g["status"] == 0)
    for i in range(N):
        r = _oracle_path(lv, i)
        assert r["status"] == 0
        assert np.array_equal(g["obs_y"][:-1, :, i], r["stop_y"])
        assert np.array_equal(g["obs_y"][-1, :, i], r["y"])
        assert g["counts"][i, 0] == r["n_accepted"]
        loss = ((np.vstack([r["stop_y"], r["y"][None]]) - lv["targets"][:, :, i]) ** 2).sum()
        assert abs(g["loss"][i] - loss) <= 1e-12 * max(1.0, loss)


def _frozen_loss(w, i, r, params, y0):
    """Loss of the frozen step sequence of path r replayed from (params, y0)."""
    t_before = np.concatenate([[0.0], r["path_t"][:-1]])
    rc, ys = O.rk_replay("lotka_volterra", y0, t_before, r["path_dt"], params=params,
                         profile=O.PROFILE_DEVICE)
    assert rc == 0
    obs = w["obs"]
    loss = 0.0
    for k, tk in enumerate(obs[:-1]):
        idx = np.nonzero(r["path_t"] == tk)[0]
        assert idx.size == 1
        loss += ((ys[idx[0]] - w["targets"][k, :, i]) ** 2).sum()
    loss += ((ys[-1] - w["targets"][-1, :, i]) ** 2).sum()
    return loss


def test_vjp_vs_frozen_step_finite_differences(gpu, orc, lv):
    g = _gpu(lv)
    h = 1e-6
    for i in range(N):
        r = _oracle_path(lv, i)
        p0, y00 = lv["params"][:, i].copy(), lv["y0"][:, i].copy()
        base = _frozen_loss(lv, i, r, p0, y00)
        assert abs(base - g["loss"][i]) <= 1e-12 * max(1.0, base)
        fd = np.zeros(6)
        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            lp = _frozen_loss(lv, i, r, p0 + e[:4], y00 + e[4:])
            lm = _frozen_loss(lv, i, r, p0 - e[:4], y00 - e[4:])
            fd[k] = (lp - lm) / (2 * h)
        adj = np.concatenate([g["grad_params"][:, i], g["grad_y0"][:, i]])
        rel = np.linalg.norm(adj - fd) / np.linalg.norm(fd)
        assert rel < 1e-6, (i, rel, adj, fd)


def test_vjp_vs_forward_sensitivity(gpu, orc, lv):
    g = _gpu(lv)
    for i in range(N):
        z0 = np.zeros(14)
        z0[:2] = lv["y0"][:, i]
        z0[10] = 1.0  # dx/dx0
        z0[13] = 1.0  # dy/dy0
        r = O.solve(O.RK45, "lv_sens", z0, 0.0, 15.0, 1e-12, 1e-14, 0.1,
                    params=lv["params"][:, i], semantics=O.INTENDED, stops=lv["obs"][:-1])
        assert r["status"] == 0
        zs = np.vstack([r["stop_y"], r["y"][None]])  # [n_obs, 14]
        grad = np.zeros(6)
        for k in range(zs.shape[0]):
            res = zs[k, :2] - lv["targets"][k, :, i]
            S = zs[k, 2:].reshape(6, 2).T  # S[:, col]
            grad += 2.0 * res @ S
        adj = np.concatenate([g["grad_params"][:, i], g["grad_y0"][:, i]])
        rel = np.linalg.norm(adj - grad) / np.linalg.norm(grad)
        assert rel < 1e-4, (i, rel)


def test_autograd_function(gpu, lv):
    import torch
    from bacon_amd.adjoint import RK45Loss
    params = _t(lv["params"]).requires_grad_(True)
    y0 = _t(lv["y0"]).requires_grad_(True)
    loss = RK45Loss.apply(params, y0, "lotka_volterra", _t(lv["obs"]), _t(lv["targets"]),
                          15.0, TOL)
    weights = torch.arange(1, N + 1, dtype=torch.float64, device="cuda:0")
    (loss * weights).sum().backward()
    g = _gpu(lv)
    w = np.arange(1, N + 1)
    assert np.allclose(params.grad.cpu().numpy(), g["grad_params"] * w, rtol=1e-14, atol=0)
    assert np.allclose(y0.grad.cpu().numpy(), g["grad_y0"] * w, rtol=1e-14, atol=0)


def test_checkpoint_capacity_exceeded(gpu, lv):
    g = _gpu(lv, max_checkpoints=2)
    assert np.all(g["status"] == 9)
    assert np.all(np.isnan(g["grad_params"]))


def test_tape_in_lds_and_hbm_agree(gpu, lv):
    """K = 16 keeps the backward pass's tape in LDS (256 lanes x 16 steps x
    4 doubles = 128 KiB); K = 48 does not fit and takes the HBM tape; K = 5
    gives short, ragged segments.  The tape holds the forward's own steps in
    every case, so the gradients are bit-identical."""
    base = _gpu(lv)
    for k, cap in ((5, 512), (48, 128)):
        g = _gpu(lv, checkpoint_every=k, max_checkpoints=cap)
        assert np.all(g["status"] == 0)
        for key in ("grad_params", "grad_y0", "loss", "obs_y"):
            assert np.array_equal(g[key], base[key]), (k, key)


def test_autotune_plan_bit_identical(gpu, lv):
    """bg_adjoint_autotune_device (bacon_amd.adjoint.autotune_vjp) times the
    gradient solve over K in {16, 32, 48} and 0-2 blocks per CU: 9 distinct
    trials from the default K = 16 in input order, and the same 9 with the
    forward pass in the order of the cost model fitted to the first trial's
    attempts, each writing valid results.  Its plan keeps K x
    max_checkpoints at the caller's 16 x 128 steps, and the solve with the
    plan (bg_adjoint_solve_tuned_device) returns the default solve's bits,
    with the fitted order switched on or off."""
    from bacon_amd.adjoint import autotune_vjp
    base = _gpu(lv)
    tn = autotune_vjp("lotka_volterra", _t(lv["y0"]), _t(lv["params"]), _t(lv["obs"]),
                      _t(lv["targets"]), 15.0, TOL)
    print("adjoint plan:", tn.as_dict())
    assert tn.n_trials == 18
    assert tn.checkpoint_every in (16, 32, 48) and tn.blocks_per_cu in (0, 1, 2)
    assert tn.max_checkpoints == -(-16 * 128 // tn.checkpoint_every)
    assert 0.0 < tn.best_ms <= tn.default_ms
    assert tn.schedule in (0, 1) and tn.forward.n_params == 4 and tn.forward.pilot_n == N
    for sched in (tn.schedule, 1 - tn.schedule):
        tn.schedule = sched
        g = _gpu(lv, plan=tn)
        assert np.all(g["status"] == 0)
        for key in ("grad_params", "grad_y0", "loss", "obs_y", "counts"):
            assert np.array_equal(g[key], base[key]), (sched, key)


def test_autotune_keeps_callers_order(gpu, lv):
    """With the caller's forward order the tuner times only K and blocks per
    CU (9 trials) and its plan has no schedule of its own."""
    import torch
    from bacon_amd.adjoint import autotune_vjp
    order = torch.arange(N - 1, -1, -1, dtype=torch.int32, device="cuda:0")
    tn = autotune_vjp("lotka_volterra", _t(lv["y0"]), _t(lv["params"]), _t(lv["obs"]),
                      _t(lv["targets"]), 15.0, TOL, order=order)
    assert tn.n_trials == 9 and tn.schedule == 0


def test_adjoint_in_measured_order_bit_identical(gpu, lv):
    """The forward pass scheduled longest first by an earlier solve's counts
    (bacon_amd.adjoint.longest_first) changes no bits."""
    from bacon_amd.adjoint import longest_first, solve_vjp
    import torch
    base = _gpu(lv)
    first = solve_vjp("lotka_volterra", _t(lv["y0"]), _t(lv["params"]), _t(lv["obs"]),
                      _t(lv["targets"]), 15.0, TOL)
    order = longest_first(first["counts"])
    torch.cuda.synchronize()
    att = first["counts"].cpu().numpy().astype(np.int64).sum(1)
    assert np.array_equal(order.cpu().numpy(), np.argsort(-att, kind="stable"))
    g = _gpu(lv, order=order)
    for key in ("grad_params", "grad_y0", "loss", "obs_y", "counts", "status"):
        assert np.array_equal(g[key], base[key]), key


def test_schedule_carried_across_solves(gpu, lv):
    """A Schedule takes the longest-first order from each solve's counts and
    hands it to the next solve, through solve_vjp and through RK45Loss; the
    results stay bit-identical to unscheduled solves."""
    import torch
    from bacon_amd.adjoint import RK45Loss, Schedule
    base = _gpu(lv)
    sch = Schedule()
    first = _gpu(lv, schedule=sch)
    assert sch.order is not None and sch.order.numel() == N
    att = first["counts"].astype(np.int64).sum(1)
    assert np.array_equal(sch.order.cpu().numpy(), np.argsort(-att, kind="stable"))
    second = _gpu(lv, schedule=sch)
    for key in ("grad_params", "grad_y0", "loss", "obs_y", "counts"):
        assert np.array_equal(first[key], base[key]), key
        assert np.array_equal(second[key], base[key]), key
    params = _t(lv["params"]).requires_grad_(True)
    y0 = _t(lv["y0"]).requires_grad_(True)
    loss = RK45Loss.apply(params, y0, "lotka_volterra", _t(lv["obs"]), _t(lv["targets"]),
                          15.0, TOL, sch)
    loss.sum().backward()
    assert np.array_equal(params.grad.cpu().numpy(), base["grad_params"])
    assert np.array_equal(y0.grad.cpu().numpy(), base["grad_y0"])


def test_forward_vs_bacon_profile(gpu, orc):
    """The GPU forward pass against the reference's own arithmetic (the
    oracle's BACON profile: separate multiply and add, err = |E|/dt compared
    with tol, delta = (tol/err)^(1/4) by pow) at all 20 observation times of
    256 instances: within 10x tol everywhere.  Lotka-Volterra orbits are
    periodic, so differences between two tol-controlled solves do not grow
    exponentially over [0, 15]."""
    import torch
    from bacon_amd.adjoint import solve_vjp
    n = 256
    w = W.lotka_volterra(n)
    obs = w["obs"]
    zeros = torch.zeros((obs.size, 2, n), dtype=torch.float64, device="cuda:0")
    g = solve_vjp("lotka_volterra", _t(w["y0"]), _t(w["params"]), _t(obs), zeros, 15.0, TOL)
    assert torch.all(g["status"] == 0)
    obs_y = g["obs_y"].cpu().numpy()
    worst = 0.0
    for i in range(n):
        r = O.solve(O.RK45, "lotka_volterra", w["y0"][:, i], 0.0, 15.0, TOL, 1e-12, 0.1,
                    params=w["params"][:, i], semantics=O.INTENDED, profile=O.PROFILE_BACON,
                    stops=obs[:-1])
        assert r["status"] == 0
        ref = np.vstack([r["stop_y"], r["y"][None]])
        worst = max(worst, np.abs(obs_y[:, :, i] - ref).max())
    assert worst <= 10 * TOL, worst


def test_observation_next_to_accepted_step(gpu, orc):
    """An observation time dt_min/8 after an accepted step of the forward
    pass: the landing keeps the controller's step size, so the solve
    finishes (a cut-down dt would fall below dt_min), the forward states are
    bit-identical to the oracle's, and the gradient is finite."""
    import torch
    from bacon_amd.adjoint import solve_vjp
    w = W.lotka_volterra(1)
    dt_min = 1e-12
    free = O.solve(O.RK45, "lotka_volterra", w["y0"][:, 0], 0.0, 15.0, TOL, dt_min, 0.1,
                   params=w["params"][:, 0], semantics=O.INTENDED, profile=O.PROFILE_DEVICE,
                   path_cap=100000)
    obs = np.array([free["path_t"][40] + dt_min / 8, 7.5, 15.0])
    zeros = torch.zeros((obs.size, 2, 1), dtype=torch.float64, device="cuda:0")
    g = solve_vjp("lotka_volterra", _t(w["y0"]), _t(w["params"]), _t(obs), zeros, 15.0, TOL,
                  dt_min=dt_min)
    assert int(g["status"][0]) == 0
    r = O.solve(O.RK45, "lotka_volterra", w["y0"][:, 0], 0.0, 15.0, TOL, dt_min, 0.1,
                params=w["params"][:, 0], semantics=O.INTENDED, profile=O.PROFILE_DEVICE,
                stops=obs[:-1])
    assert r["status"] == 0
    obs_y = g["obs_y"].cpu().numpy()
    assert np.array_equal(obs_y[:-1, :, 0], r["stop_y"])
    assert np.array_equal(obs_y[-1, :, 0], r["y"])
    assert torch.all(torch.isfinite(g["grad_params"])) and torch.all(torch.isfinite(g["grad_y0"]))
