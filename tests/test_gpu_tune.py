"""Autotuner (bg_rk_autotune_device / bg_rk_schedule_device /
bg_rk_solve_tuned_device) on a Lorenz-63 ensemble of the batched workload.

Results of an RK solve do not depend on the schedule or launch shape, so the
tuned solve must be bit-identical to the untuned one (and to the oracle's
DEVICE profile on sampled instances); the schedule must be a permutation; the
fitted cost model must rank instances by their actual attempt counts.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from bacon_amd import _lib as L
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu

N = 65536


@pytest.fixture(scope="module")
def setup():
    import torch
    dev = "cuda:0"
    w = W.lorenz_batched(N)
    T = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    b = dict(y0=T(w["y0"]), params=T(w["params"]), tend=T(w["t_end"]), tol=T(w["tol"]))

    def desc(outs):
        return L.EnsembleDesc(
            n=N, dim=3, n_params=3, y0=b["y0"].data_ptr(), params=b["params"].data_ptr(),
            t0_s=0.0, t_end=b["tend"].data_ptr(), tol=b["tol"].data_ptr(),
            dt_min_s=w["dt_min"], dt_max_s=w["dt_max"],
            t_out=outs["t"].data_ptr(), y_out=outs["y"].data_ptr(),
            status=outs["st"].data_ptr(), counts=outs["cnt"].data_ptr())

    def outs():
        return dict(t=torch.zeros(N, dtype=torch.float64, device=dev),
                    y=torch.zeros((3, N), dtype=torch.float64, device=dev),
                    st=torch.full((N,), -7, dtype=torch.int32, device=dev),
                    cnt=torch.zeros((N, 2), dtype=torch.int32, device=dev))

    # reference: untuned solve in input order
    ref = outs()
    d = desc(ref)
    assert L.lib().bg_rk_solve_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d), None,
                                      None) == 0
    torch.cuda.synchronize()
    tuned_out = outs()
    dt = desc(tuned_out)
    tn = L.Tuning()
    rc = L.lib().bg_rk_autotune_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(dt),
                                       C.byref(tn), None)
    assert rc == 0
    return dict(w=w, b=b, desc=desc, outs=outs, ref=ref, tn=tn, tuned_out=tuned_out)


def _np(o):
    return {k: v.cpu().numpy() for k, v in o.items()}


def test_tuning_fields(setup):
    tn = setup["tn"]
    assert tn.n_params == 3 and tn.pilot_n == 16384 and tn.n_trials == 6
    assert tn.launch.block in (64, 128, 256)
    assert 0.0 < tn.best_ms <= tn.default_ms
    # log cost ~ log(t_end) - log(tol)/4 for an adaptive 4th-order method
    assert 0.8 < tn.coef[1] < 1.2 and -0.35 < tn.coef[2] < -0.15
    assert tn.coef[3] == 0.0 and tn.coef[5] == 0.0  # sigma, beta do not vary


def test_autotune_trials_write_valid_results(setup):
    a, r = _np(setup["tuned_out"]), _np(setup["ref"])
    for k in ("t", "y", "st", "cnt"):
        assert np.array_equal(a[k], r[k]), k


def test_solve_tuned_bit_identical_and_oracle(setup, orc):
    import torch
    o = setup["outs"]()
    d = setup["desc"](o)
    rc = L.lib().bg_rk_solve_tuned_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d),
                                          C.byref(setup["tn"]), None)
    assert rc == 0
    torch.cuda.synchronize()
    a, r = _np(o), _np(setup["ref"])
    for k in ("t", "y", "st", "cnt"):
        assert np.array_equal(a[k], r[k]), k
    w = setup["w"]
    for i in range(0, N, 8191):
        orr = O.solve(O.RK45, "lorenz", w["y0"][:, i], 0.0, w["t_end"][i], w["tol"][i],
                      w["dt_min"], w["dt_max"], params=w["params"][:, i],
                      semantics=O.INTENDED, profile=O.PROFILE_DEVICE)
        assert np.array_equal(a["y"][:, i], orr["y"]) and a["cnt"][i, 0] == orr["n_accepted"]


def test_schedule_is_a_longest_first_permutation(setup):
    import torch
    order = torch.full((N,), -1, dtype=torch.int32, device="cuda:0")
    d = setup["desc"](setup["outs"]())
    rc = L.lib().bg_rk_schedule_device(C.byref(setup["tn"]), C.byref(d),
                                       C.c_void_p(order.data_ptr()), None)
    assert rc == 0
    torch.cuda.synchronize()
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(N))
    att = _np(setup["ref"])["cnt"].astype(np.int64).sum(1)
    # rank correlation between schedule position and actual work
    pos = np.empty(N)
    pos[o] = np.arange(N)
    ra = np.argsort(np.argsort(-att))
    rho = np.corrcoef(pos, ra)[0, 1]
    assert rho > 0.9, rho


def test_autotune_rejects_given_order(setup):
    import torch
    d = setup["desc"](setup["outs"]())
    order = torch.arange(N, dtype=torch.int32, device="cuda:0")
    d.order = order.data_ptr()
    tn = L.Tuning()
    assert L.lib().bg_rk_autotune_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d),
                                         C.byref(tn), None) == -1


def test_schedule_by_counts_is_a_stable_descending_sort(setup):
    """bg_schedule_by_counts_device against numpy's stable sort of the same
    keys: the reference solve's attempts, plus a synthetic set with many
    ties, a saturating sum and n = 1; n = 0 writes nothing and a missing
    pointer is rejected."""
    import torch
    dev = "cuda:0"
    att = _np(setup["ref"])["cnt"]
    rng = np.random.default_rng(7)
    ties = np.stack([rng.integers(0, 4, 5000), rng.integers(0, 3, 5000)], 1)
    ties[17] = (0xFFFFFFF0, 0x20)  # sum past 2^32 - 1 counts as the maximum
    for cnt in (att, ties, att[:1]):
        n = cnt.shape[0]
        c = torch.as_tensor(cnt.astype(np.uint32).view(np.int32), device=dev)
        order = torch.full((n,), -1, dtype=torch.int32, device=dev)
        assert L.lib().bg_schedule_by_counts_device(c.data_ptr(), n, order.data_ptr(),
                                                    None) == 0
        torch.cuda.synchronize()
        key = np.minimum(cnt[:, 0].astype(np.uint64) + cnt[:, 1].astype(np.uint64),
                         0xFFFFFFFF)
        want = np.argsort(-key.astype(np.int64), kind="stable")
        assert np.array_equal(order.cpu().numpy(), want)
    order = torch.full((4,), -1, dtype=torch.int32, device=dev)
    assert L.lib().bg_schedule_by_counts_device(None, 0, order.data_ptr(), None) == 0
    assert np.all(order.cpu().numpy() == -1)
    assert L.lib().bg_schedule_by_counts_device(None, 4, order.data_ptr(), None) == -1


def test_solve_in_measured_order_bit_identical(setup):
    """A solve scheduled by the reference solve's own counts returns the
    same bits per instance."""
    import torch
    ref = _np(setup["ref"])
    c = setup["ref"]["cnt"]
    order = torch.empty(N, dtype=torch.int32, device="cuda:0")
    assert L.lib().bg_schedule_by_counts_device(c.data_ptr(), N, order.data_ptr(), None) == 0
    o = setup["outs"]()
    d = setup["desc"](o)
    d.order = order.data_ptr()
    assert L.lib().bg_rk_solve_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d), None,
                                      None) == 0
    torch.cuda.synchronize()
    got = _np(o)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k
