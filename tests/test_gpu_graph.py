"""HIP graph capture of batched solves (bg_graph_capture_bdf / _rk).

Each graph is replayed several times with the program's own eager solves and
host synchronisations in between, and every replay is compared bit for bit
with the eager solve of the same inputs (a lane's or group's arithmetic does
not depend on scheduling, so equal inputs give equal bits).  The inputs are
then rewritten in place: the replay must solve the new values.
"""
import ctypes as C

import numpy as np
import pytest

from bacon_amd import _lib as L
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu


def _bufs(w, n, dim):
    import torch
    dev = "cuda:0"
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    return dict(y0=T(w["y0"]), params=T(w["params"]),
                t=torch.zeros(n, dtype=torch.float64, device=dev),
                y=torch.zeros((dim, n), dtype=torch.float64, device=dev),
                st=torch.full((n,), -7, dtype=torch.int32, device=dev),
                cnt=torch.zeros((n, 2), dtype=torch.int32, device=dev))


def _desc(b, n, dim, npar, t_end, tol, dt_min, dt_max, t_end_arr=None, tol_arr=None):
    return L.EnsembleDesc(
        n=n, dim=dim, n_params=npar, y0=b["y0"].data_ptr(), params=b["params"].data_ptr(),
        t0_s=0.0, t_end_s=t_end, tol_s=tol, dt_min_s=dt_min, dt_max_s=dt_max,
        t_end=t_end_arr, tol=tol_arr,
        t_out=b["t"].data_ptr(), y_out=b["y"].data_ptr(), status=b["st"].data_ptr(),
        counts=b["cnt"].data_ptr())


def _snap(b):
    return {k: b[k].cpu().numpy().copy() for k in ("t", "y", "st", "cnt")}


def _clear(b):
    b["t"].zero_()
    b["y"].zero_()
    b["st"].fill_(-7)
    b["cnt"].zero_()


def _check_replays(b, d, eager, graph, new_y0):
    import torch
    lib = L.lib()
    ref = None
    for rep in range(3):
        _clear(b)
        assert eager(d) == 0
        torch.cuda.synchronize()
        ref = _snap(b)
        assert np.all(ref["st"] == 0)
        _clear(b)
        torch.cuda.synchronize()
        assert lib.bg_graph_launch(graph, None) == 0
        torch.cuda.synchronize()
        got = _snap(b)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), (rep, k)
    # new inputs in the same buffers
    b["y0"].copy_(new_y0)
    _clear(b)
    assert eager(d) == 0
    torch.cuda.synchronize()
    ref2 = _snap(b)
    assert not np.array_equal(ref2["y"], ref["y"])
    _clear(b)
    assert lib.bg_graph_launch(graph, None) == 0
    torch.cuda.synchronize()
    got = _snap(b)
    for k in ref2:
        assert np.array_equal(got[k], ref2[k]), k


def test_bdf_graph_replays(gpu):
    import torch
    n = 256
    w = W.brusselator(n)
    b = _bufs(w, n, 32)
    d = _desc(b, n, 32, 3, 10.0, 1e-7, 1e-12, 0.1)
    lib = L.lib()
    g = C.c_void_p()
    assert lib.bg_graph_capture_bdf(L.BDF6, L.RHS["brusselator"], L.INTENDED, 64, C.byref(d),
                                    C.byref(g)) == 0
    try:
        eager = lambda dd: lib.bg_bdf_solve_device(L.BDF6, L.RHS["brusselator"], L.INTENDED, 64,
                                                   C.byref(dd), None)
        w2 = W.brusselator(n, first=n)
        _check_replays(b, d, eager, g, torch.as_tensor(w2["y0"], device="cuda:0"))
    finally:
        lib.bg_graph_destroy(g)


def test_bdf_fp32_graph_replays(gpu):
    """The FP32 Jacobian inverse under capture: its float pivot-row buffer
    and the graph-owned scratch of shared finite-difference columns give
    replays bit-identical to eager FP32 solves."""
    import torch
    n = 256
    w = W.brusselator(n)
    b = _bufs(w, n, 32)
    d = _desc(b, n, 32, 3, 10.0, 1e-7, 1e-12, 0.1)
    lib = L.lib()
    g = C.c_void_p()
    assert lib.bg_graph_capture_bdf(L.BDF6, L.RHS["brusselator"], L.INTENDED, 32, C.byref(d),
                                    C.byref(g)) == 0
    try:
        eager = lambda dd: lib.bg_bdf_solve_device(L.BDF6, L.RHS["brusselator"], L.INTENDED, 32,
                                                   C.byref(dd), None)
        w2 = W.brusselator(n, first=n)
        _check_replays(b, d, eager, g, torch.as_tensor(w2["y0"], device="cuda:0"))
    finally:
        lib.bg_graph_destroy(g)


def test_rk_graph_replays(gpu):
    import torch
    n = 4096
    w = W.lorenz_plain(n)
    b = _bufs(w, n, 3)
    d = _desc(b, n, 3, 3, 10.0, 1e-6, 1e-6, 0.1)
    lib = L.lib()
    g = C.c_void_p()
    lc = L.Launch(block=256, blocks_per_cu=2)
    assert lib.bg_graph_capture_rk(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d), C.byref(lc),
                                   C.byref(g)) == 0
    try:
        eager = lambda dd: lib.bg_rk_solve_device(L.RK45, L.RHS["lorenz"], L.INTENDED,
                                                  C.byref(dd), None, None)
        w2 = W.lorenz_plain(n, first=n)
        _check_replays(b, d, eager, g, torch.as_tensor(w2["y0"], device="cuda:0"))
    finally:
        lib.bg_graph_destroy(g)


def test_graph_capture_rejects_bad_arguments(gpu):
    lib = L.lib()
    g = C.c_void_p()
    d = L.EnsembleDesc(n=0, dim=3, n_params=3)
    assert lib.bg_graph_capture_rk(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d), None,
                                   C.byref(g)) == -1
    assert lib.bg_graph_capture_bdf(L.RK45, L.RHS["brusselator"], L.INTENDED, 64, C.byref(d),
                                    C.byref(g)) == -1
    assert lib.bg_graph_launch(None, None) == -1
