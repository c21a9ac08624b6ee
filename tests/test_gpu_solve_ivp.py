"""Batched solve_ivp (bg_ensemble_solve_ivp): per instance Adams5, then RK45
for the instances that did not finish OK, then BDF6 (README.md:45-47).
Each instance's final result must equal the direct ensemble solve of the
method that produced it, bit for bit."""
import numpy as np
import pytest

from bacon_amd import _lib as L
from bacon_amd import ivp
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu


def _ensemble(method, w, max_attempts=0):
    n = w["y0"].shape[1]
    e = (ivp.Ensemble(method, w["dim"], n)
         .with_minimum_dt(w["dt_min"]).with_maximum_dt(w["dt_max"])
         .with_tolerance(w["tol"]).with_initial_time(w.get("t0", 0.0))
         .with_ending_time(w["t_end"]).with_initial_conditions(w["y0"])
         .with_derivative(w["rhs"], w["params"] if w["params"].size else None)
         .with_semantics(L.INTENDED))
    if max_attempts:
        e.with_max_attempts(max_attempts)
    return e


def _same(a, b, i):
    return (a["t"][i] == b["t"][i] and np.array_equal(a["y"][:, i], b["y"][:, i])
            and a["status"][i] == b["status"][i] and a["n_accepted"][i] == b["n_accepted"][i])


def test_adams_then_rk45(gpu):
    """Lorenz-63 with a 5000-step cap: Adams5 finishes some instances and
    hits the cap on others, which RK45 then solves."""
    w = W.lorenz_plain(128)
    r = _ensemble(L.RK45, w, 5000).solve_ivp()
    assert np.all(r["status"] == 0)
    used = r["method_used"]
    assert set(np.unique(used)) == {L.ADAMS5, L.RK45}
    a = _ensemble(L.ADAMS5, w, 5000).solve()
    k = _ensemble(L.RK45, w, 5000).solve()
    for i in range(128):
        ref = a if used[i] == L.ADAMS5 else k
        assert _same(r, ref, i), i
        if used[i] == L.RK45:
            assert a["status"][i] == 9


def test_bdf6_for_stiff_brusselator(gpu):
    """The 32-component Brusselator has BDF kernels only: the chain skips
    Adams5 and RK45 and solves every instance with BDF6."""
    w = W.brusselator(64)
    r = _ensemble(L.RK45, w).solve_ivp()
    assert np.all(r["method_used"] == L.BDF6)
    b = _ensemble(L.BDF6, w).solve()
    for i in range(64):
        assert _same(r, b, i), i


def test_adams_alone_when_it_succeeds(gpu):
    n = 32
    w = dict(rhs="exp", dim=1, y0=np.ones((1, n)), params=np.zeros((0, n)), t0=0.0,
             t_end=np.linspace(1.0, 2.0, n), tol=5e-4, dt_min=1e-5, dt_max=0.1)
    r = _ensemble(L.RK45, w).solve_ivp()
    assert np.all(r["method_used"] == L.ADAMS5) and np.all(r["status"] == 0)
    a = _ensemble(L.ADAMS5, w).solve()
    for i in range(n):
        assert _same(r, a, i)
    assert np.all(np.abs(r["y"][0] - np.exp(w["t_end"])) <= 0.01 * np.exp(w["t_end"]))
