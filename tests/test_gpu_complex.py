"""Complex state on the GPU (the reference's N: ComplexField, rk.rs:53,
adams.rs:44): the two-level system c' = (-i H - gamma) c stored as (re, im)
pairs (BG_RHS_TWO_LEVEL, dim 4).

  * RK45, RK23 and Adams5 ensembles: bit-identical to the oracle's DEVICE
    profile (t, y, status, step counts) on 256 instances;
  * within 10x tol of the BACON profile, whose error norm groups each
    complex component as nalgebra's dotc does;
  * within 2e-6 of the closed form e^{-gamma t} exp(-iHt) c0 at t = 10;
  * BDF refuses complex state (its Broyden update uses the plain transpose,
    bdf.rs:461-466, so the complex run is not the real-pair run).
"""
import numpy as np
import pytest

import oracle as O
from bacon_amd import _lib as L
from bacon_amd import ivp
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu

N = 256
METHODS = [(L.RK45, O.RK45), (L.RK23, O.RK23), (L.ADAMS5, O.ADAMS5)]


@pytest.fixture(scope="module")
def w():
    return W.two_level(N)


@pytest.mark.parametrize("gm,om", METHODS, ids=["rk45", "rk23", "adams5"])
def test_two_level_vs_oracle(gpu, orc, w, gm, om):
    g = ivp.solve_ensemble(gm, w)
    ex = W.two_level_exact(w["params"], w["y0"], 10.0)
    assert np.all(g["status"] == 0) and np.all(g["t"] == 10.0)
    worst_bacon = 0.0
    for i in range(N):
        kw = dict(params=w["params"][:, i], semantics=O.INTENDED)
        d = O.solve(om, "two_level", w["y0"][:, i], 0.0, 10.0, 1e-8, 1e-12, 0.1,
                    profile=O.PROFILE_DEVICE, **kw)
        assert np.array_equal(g["y"][:, i], d["y"]), i
        assert g["n_accepted"][i] == d["n_accepted"] and g["n_rejected"][i] == d["n_rejected"]
        b = O.solve(om, "two_level", w["y0"][:, i], 0.0, 10.0, 1e-8, 1e-12, 0.1,
                    profile=O.PROFILE_BACON, **kw)
        assert b["status"] == 0
        worst_bacon = max(worst_bacon, np.abs(g["y"][:, i] - b["y"]).max() / 1e-8)
    assert worst_bacon < 10.0, worst_bacon
    assert np.abs(g["y"] - ex).max() < 2e-6
    print(f"max |GPU - BACON| / tol = {worst_bacon:.3g}, "
          f"max |GPU - exact| = {np.abs(g['y'] - ex).max():.3g}")


def test_bdf_refuses_complex_state(gpu, w):
    import ctypes as C
    import torch
    n = 8
    y0 = torch.as_tensor(np.ascontiguousarray(w["y0"][:, :n]), device="cuda:0")
    pr = torch.as_tensor(np.ascontiguousarray(w["params"][:, :n]), device="cuda:0")
    out = dict(t=torch.zeros(n, dtype=torch.float64, device="cuda:0"),
               y=torch.zeros((4, n), dtype=torch.float64, device="cuda:0"),
               st=torch.zeros(n, dtype=torch.int32, device="cuda:0"))
    d = L.EnsembleDesc(n=n, dim=4, n_params=3, y0=y0.data_ptr(), params=pr.data_ptr(),
                       t0_s=0.0, t_end_s=1.0, tol_s=1e-8, dt_min_s=1e-12, dt_max_s=0.1,
                       t_out=out["t"].data_ptr(), y_out=out["y"].data_ptr(),
                       status=out["st"].data_ptr())
    assert L.lib().bg_bdf_solve_device(L.BDF6, L.RHS["two_level"], L.INTENDED, 64, C.byref(d),
                                       None) == L.ERR_UNSUPPORTED
