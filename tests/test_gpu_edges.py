"""Edge cases of the device entry points, for every method family.

  * n = 0: the call returns OK and writes nothing;
  * instances that start at their end time (t0 == t_end) finish at once:
    status OK, y = y0, t = t0, no steps;
  * results do not depend on the batch around an instance: a single
    instance, and a group one past a full wave or MFMA tile, reproduce the
    same instances of a larger batch bit for bit (each instance's arithmetic
    is its own lane's, group's or tile column's);
  * descriptors with a missing input are rejected with -1
    (BG_ERR_INVALID_ARGUMENT), a dimension no kernel serves with -1 or -3
    (BG_ERR_UNSUPPORTED), before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

from bacon_amd import _lib as L
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
SENTINEL = -7


def _t(a, dtype=None):
    import torch
    return torch.as_tensor(np.array(a), dtype=dtype or torch.float64, device=DEV)


def _bufs(w, n, t_end=None):
    """Device inputs and sentinel-filled outputs for the first n instances."""
    import torch
    dim = w["dim"]
    params = w["params"][:, :n] if w["params"].size else np.zeros((1, n))
    te = np.broadcast_to(np.asarray(w["t_end"], dtype=np.float64), (w["y0"].shape[1],))[:n]
    tl = np.broadcast_to(np.asarray(w["tol"], dtype=np.float64), (w["y0"].shape[1],))[:n]
    b = dict(y0=_t(w["y0"][:, :n].reshape(dim, n)), params=_t(params.reshape(-1, n)),
             t_end=_t(te if t_end is None else t_end), tol=_t(tl),
             t=torch.full((n,), float(SENTINEL), dtype=torch.float64, device=DEV),
             y=torch.full((dim, n), float(SENTINEL), dtype=torch.float64, device=DEV),
             st=torch.full((n,), SENTINEL, dtype=torch.int32, device=DEV),
             cnt=torch.full((n, 2), SENTINEL, dtype=torch.int32, device=DEV))
    if w.get("shared") is not None:
        b["shared"] = _t(w["shared"])
    return b


def _desc(w, b, n):
    return L.EnsembleDesc(
        n=n, dim=w["dim"], n_params=b["params"].shape[0],
        y0=b["y0"].data_ptr() if n else None, params=b["params"].data_ptr() if n else None,
        shared=b["shared"].data_ptr() if "shared" in b else None,
        t0_s=w.get("t0", 0.0), t_end=b["t_end"].data_ptr() if n else None,
        tol=b["tol"].data_ptr() if n else None, dt_min_s=w["dt_min"], dt_max_s=w["dt_max"],
        t_out=b["t"].data_ptr() if n else None, y_out=b["y"].data_ptr() if n else None,
        status=b["st"].data_ptr() if n else None, counts=b["cnt"].data_ptr() if n else None)


def _host(b):
    import torch
    torch.cuda.synchronize()
    return {k: b[k].cpu().numpy() for k in ("t", "y", "st", "cnt")}


# (name, workload factory, solve(desc) -> rc)
def _rk(d):
    return L.lib().bg_rk_solve_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d), None, None)


def _adams(d):
    return L.lib().bg_adams_solve_device(L.ADAMS5, L.RHS["lorenz"], L.INTENDED, C.byref(d), None)


def _bdf(d):
    return L.lib().bg_bdf_solve_device(L.BDF6, L.RHS["brusselator"], L.INTENDED, 64,
                                       C.byref(d), None)


def _linear(d):
    return L.lib().bg_rk_solve_device(L.RK45, L.RHS["linear"], L.INTENDED, C.byref(d), None,
                                      None)


def _lorenz(n):
    return W.lorenz_plain(n)


def _bruss(n):
    w = W.brusselator(n)
    w["t_end"] = 2.0
    return w


FAMILIES = [
    ("rk45", _lorenz, _rk, 65),
    ("adams5", _lorenz, _adams, 65),
    ("bdf6", _bruss, _bdf, 5),
    ("linear", W.linear_shared, _linear, 17),
]


@pytest.mark.parametrize("name,make,solve,ragged", FAMILIES, ids=[f[0] for f in FAMILIES])
def test_empty_batch_writes_nothing(gpu, name, make, solve, ragged):
    w = make(4)
    b = _bufs(w, 4)
    d = _desc(w, b, 4)
    d.n = 0
    assert solve(d) == 0
    h = _host(b)
    assert np.all(h["st"] == SENTINEL) and np.all(h["t"] == SENTINEL)
    assert np.all(h["y"] == SENTINEL) and np.all(h["cnt"] == SENTINEL)


@pytest.mark.parametrize("name,make,solve,ragged", FAMILIES, ids=[f[0] for f in FAMILIES])
def test_instances_starting_at_their_end(gpu, name, make, solve, ragged):
    n = 6
    w = make(n)
    t0 = w.get("t0", 0.0)
    te = np.broadcast_to(np.asarray(w["t_end"], dtype=np.float64), (w["y0"].shape[1],))[:n].copy()
    te[::2] = t0  # every other instance is done before it starts
    b = _bufs(w, n, t_end=te)
    assert solve(_desc(w, b, n)) == 0
    h = _host(b)
    assert np.all(h["st"] == 0)
    done = np.arange(n) % 2 == 0
    assert np.all(h["t"][done] == t0)
    assert np.array_equal(h["y"][:, done], w["y0"][:, :n][:, done])
    assert np.all(h["cnt"][done] == 0)
    # RK and Adams land on t_end exactly; BDF's time sum ends within rounding
    assert np.allclose(h["t"][~done], te[~done], rtol=1e-9, atol=0)
    assert np.all(h["cnt"][~done, 0] > 0)


@pytest.mark.parametrize("name,make,solve,ragged", FAMILIES, ids=[f[0] for f in FAMILIES])
def test_results_independent_of_the_batch(gpu, name, make, solve, ragged):
    big = 4 * ragged
    w = make(big)
    ref = _bufs(w, big)
    assert solve(_desc(w, ref, big)) == 0
    r = _host(ref)
    for n in (1, ragged):
        b = _bufs(w, n)
        assert solve(_desc(w, b, n)) == 0
        h = _host(b)
        assert np.array_equal(h["st"], r["st"][:n]), (name, n)
        assert np.array_equal(h["t"], r["t"][:n]), (name, n)
        assert np.array_equal(h["y"], r["y"][:, :n]), (name, n)
        assert np.array_equal(h["cnt"], r["cnt"][:n]), (name, n)


@pytest.mark.parametrize("name,make,solve,ragged", FAMILIES, ids=[f[0] for f in FAMILIES])
def test_invalid_descriptors_rejected(gpu, name, make, solve, ragged):
    w = make(4)
    b = _bufs(w, 4)
    d = _desc(w, b, 4)
    d.y0 = None
    assert solve(d) == -1
    d = _desc(w, b, 4)
    d.dim = w["dim"] + 1
    assert solve(d) in (-1, -3)  # invalid argument, or no kernel for that dimension
    h = _host(b)
    assert np.all(h["st"] == SENTINEL)


def test_adjoint_single_and_empty(gpu):
    """The discrete adjoint of one instance equals that instance's gradient
    in a batch of 9 (LDS tape, one instance per lane); n = 0 is a no-op."""
    import torch
    from bacon_amd.adjoint import solve_vjp
    w = W.lotka_volterra(9)
    obs = _t(w["obs"])
    zeros = torch.zeros((w["obs"].size, 2, 9), dtype=torch.float64, device=DEV)
    tgt = solve_vjp("lotka_volterra", _t(w["y0"]), _t(w["target_params"]), obs, zeros,
                    15.0, 1e-8)["obs_y"]
    full = solve_vjp("lotka_volterra", _t(w["y0"]), _t(w["params"]), obs, tgt, 15.0, 1e-8)
    one = solve_vjp("lotka_volterra", _t(w["y0"][:, :1]), _t(w["params"][:, :1]), obs,
                    tgt[:, :, :1].contiguous(), 15.0, 1e-8)
    for k in ("loss", "grad_params", "grad_y0", "status", "counts"):
        assert torch.equal(one[k], full[k][..., :1] if full[k].dim() > 1 and k != "counts"
                           else full[k][:1]), k
    d = L.AdjointDesc(base=L.EnsembleDesc(n=0, dim=2, n_params=4), n_obs=0)
    assert L.lib().bg_rk_vjp_device(L.RK45, L.RHS["lotka_volterra"], C.byref(d), None) == 0
