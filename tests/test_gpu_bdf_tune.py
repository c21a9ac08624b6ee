"""BDF autotuner (bg_bdf_autotune_device / bg_bdf_solve_tuned_device) on
Brusselator ensembles (32 components: one instance per wave) and on the
Robertson kinetics (3 components: lane groups).

A BDF instance's arithmetic does not depend on the schedule or on how many
waves are resident, so every trial of the tuner and the tuned solve must be
bit-identical to the untuned solve in input order (t, y, status, step counts
and the corrector's work counters); the fitted schedule must be a
permutation; the plan's fields must describe the trials that ran.
"""
import ctypes as C

import numpy as np
import pytest

from bacon_amd import _lib as L
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = {
    # name: (workload, n, t_end, method, lu_bits)
    "brusselator_bdf6": (W.brusselator, 8192, 2.0, L.BDF6, 64),
    "brusselator_bdf6_lu32": (W.brusselator, 4096, 2.0, L.BDF6, 32),
    "robertson_bdf2": (W.robertson, 8192, 1.0, L.BDF2, 64),
}


def _setup(name):
    import torch
    dev = "cuda:0"
    fn, n, t_end, method, bits = CASES[name]
    w = fn(n)
    dim = w["dim"]
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    b = dict(y0=T(w["y0"]), params=T(w["params"]))

    def outs():
        return dict(t=torch.zeros(n, dtype=torch.float64, device=dev),
                    y=torch.zeros((dim, n), dtype=torch.float64, device=dev),
                    st=torch.full((n,), -7, dtype=torch.int32, device=dev),
                    cnt=torch.zeros((n, 2), dtype=torch.int32, device=dev),
                    work=torch.zeros((n, 2), dtype=torch.int32, device=dev))

    def desc(o):
        return L.EnsembleDesc(
            n=n, dim=dim, n_params=w["params"].shape[0], y0=b["y0"].data_ptr(),
            params=b["params"].data_ptr(), t0_s=0.0, t_end_s=t_end, tol_s=float(w["tol"][0]),
            dt_min_s=w["dt_min"], dt_max_s=w["dt_max"], t_out=o["t"].data_ptr(),
            y_out=o["y"].data_ptr(), status=o["st"].data_ptr(), counts=o["cnt"].data_ptr(),
            work=o["work"].data_ptr())

    rhs = L.RHS[w["rhs"]]
    ref = outs()
    assert L.lib().bg_bdf_solve_device(method, rhs, L.INTENDED, bits, C.byref(desc(ref)),
                                       None) == 0
    torch.cuda.synchronize()
    tuned = outs()
    tn = L.Tuning()
    rc = L.lib().bg_bdf_autotune_device(method, rhs, L.INTENDED, bits, C.byref(desc(tuned)),
                                        C.byref(tn), None)
    assert rc == 0
    return dict(n=n, method=method, bits=bits, rhs=rhs, outs=outs, desc=desc, ref=ref,
                tuned=tuned, tn=tn)


@pytest.fixture(scope="module", params=list(CASES))
def setup(request):
    return _setup(request.param)


def _np(o):
    return {k: v.cpu().numpy() for k, v in o.items()}


def _same(a, b):
    a, b = _np(a), _np(b)
    for k in b:
        assert np.array_equal(a[k], b[k]), k


def test_plan_fields(setup):
    tn, n = setup["tn"], setup["n"]
    assert tn.launch.block == 64
    assert tn.launch.blocks_per_cu in (0, 16, 12, 8, 4)
    assert 1 <= tn.n_trials <= 5
    assert tn.pilot_n == min(n, 16384) and tn.n_params == 3
    assert 0.0 < tn.best_ms <= tn.default_ms and tn.pilot_ms > 0.0
    print(f"plan: cap {tn.launch.blocks_per_cu}, pilot {tn.pilot_ms:.2f} ms, default "
          f"{tn.default_ms:.2f} ms, best {tn.best_ms:.2f} ms, {tn.n_trials} trials")


def test_trials_bit_identical(setup):
    """The tuner's last trial left its solution in the caller's outputs."""
    assert np.all(_np(setup["ref"])["st"] == 0)
    _same(setup["tuned"], setup["ref"])


def test_solve_tuned_bit_identical(setup):
    import torch
    o = setup["outs"]()
    rc = L.lib().bg_bdf_solve_tuned_device(setup["method"], setup["rhs"], L.INTENDED,
                                           setup["bits"], C.byref(setup["desc"](o)),
                                           C.byref(setup["tn"]), None)
    assert rc == 0
    torch.cuda.synchronize()
    _same(o, setup["ref"])


def test_every_cap_bit_identical(setup):
    """The shapes the tuner tries, forced one by one through a plan."""
    import torch
    for cap in (4, 12):
        tn = L.Tuning()
        C.memmove(C.byref(tn), C.byref(setup["tn"]), C.sizeof(L.Tuning))
        tn.launch.blocks_per_cu = cap
        o = setup["outs"]()
        assert L.lib().bg_bdf_solve_tuned_device(setup["method"], setup["rhs"], L.INTENDED,
                                                 setup["bits"], C.byref(setup["desc"](o)),
                                                 C.byref(tn), None) == 0
        torch.cuda.synchronize()
        _same(o, setup["ref"])


def test_schedule_is_a_permutation_ranked_by_work(setup):
    import torch
    n = setup["n"]
    order = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
    d = setup["desc"](setup["outs"]())
    assert L.lib().bg_rk_schedule_device(C.byref(setup["tn"]), C.byref(d),
                                         C.c_void_p(order.data_ptr()), None) == 0
    torch.cuda.synchronize()
    o = order.cpu().numpy()
    assert np.array_equal(np.sort(o), np.arange(n))
    wk = _np(setup["ref"])["work"].astype(np.int64)
    cost = 16 * wk[:, 1] + wk[:, 0]
    pos = np.empty(n)
    pos[o] = np.arange(n)
    rho = np.corrcoef(pos, np.argsort(np.argsort(-cost)))[0, 1]
    print(f"rank correlation of schedule position and corrector work: {rho:.3f}")
    assert rho > 0.0


def test_rejects_given_order_and_non_bdf():
    import torch
    s = _setup("robertson_bdf2")
    d = s["desc"](s["outs"]())
    order = torch.arange(s["n"], dtype=torch.int32, device="cuda:0")
    d.order = order.data_ptr()
    tn = L.Tuning()
    lib = L.lib()
    assert lib.bg_bdf_autotune_device(L.BDF2, s["rhs"], L.INTENDED, 64, C.byref(d),
                                      C.byref(tn), None) == L.ERR_INVALID_ARGUMENT
    d.order = None
    assert lib.bg_bdf_autotune_device(L.RK45, s["rhs"], L.INTENDED, 64, C.byref(d),
                                      C.byref(tn), None) == L.ERR_INVALID_ARGUMENT
    assert lib.bg_bdf_autotune_device(L.BDF2, s["rhs"], L.INTENDED, 16, C.byref(d),
                                      C.byref(tn), None) == L.ERR_INVALID_ARGUMENT
