"""bg_rk_solve_device_noalloc (through the C ABI) against bg_rk_solve_device.

The allocation-free entry point runs the same kernels with the launch
workspace in the library's static device array, so its results must equal
the allocating entry point's bit for bit.  Two calls in a row share that
array; the second, smaller one must not see the first one's counter or
descriptor.  Both the one-lane-per-instance RK kernels (Lorenz-63, RK45 and
RK23) and the shared-matrix MFMA kernel (linear, N = 128) are covered.
"""
import ctypes as C

import numpy as np
import pytest

from bacon_amd import _lib as L
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu


def _buffers(torch, w, n):
    dev = "cuda:0"
    dim = w["dim"]

    def T(a, dt=torch.float64):
        return torch.as_tensor(np.array(a, copy=True), dtype=dt, device=dev)

    P = w["params"].shape[0]
    b = dict(y0=T(w["y0"][:, :n]), t_end=T(np.broadcast_to(w["t_end"], (w["y0"].shape[1],))[:n]),
             tol=T(np.broadcast_to(w["tol"], (w["y0"].shape[1],))[:n]))
    b["params"] = T(w["params"][:, :n]) if P else None
    b["shared"] = T(w["shared"]) if w.get("shared") is not None else None
    return b


def _solve(torch, entry, method, w, b, n):
    dev = "cuda:0"
    dim, P = w["dim"], w["params"].shape[0]
    out = dict(t=torch.full((n,), -1.0, dtype=torch.float64, device=dev),
               y=torch.full((dim, n), -1.0, dtype=torch.float64, device=dev),
               st=torch.full((n,), -7, dtype=torch.int32, device=dev),
               cnt=torch.zeros((n, 2), dtype=torch.int32, device=dev))
    d = L.EnsembleDesc(
        n=n, dim=dim, n_params=P, y0=b["y0"].data_ptr(),
        params=b["params"].data_ptr() if b["params"] is not None else None,
        shared=b["shared"].data_ptr() if b["shared"] is not None else None,
        t0_s=0.0, t_end=b["t_end"].data_ptr(), tol=b["tol"].data_ptr(),
        dt_min_s=w["dt_min"], dt_max_s=w["dt_max"],
        t_out=out["t"].data_ptr(), y_out=out["y"].data_ptr(), status=out["st"].data_ptr(),
        counts=out["cnt"].data_ptr())
    sh = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(L.lib(), entry)(method, L.RHS[w["rhs"]], L.INTENDED, C.byref(d), None, sh)
    assert rc == 0, (entry, rc)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


CASES = [
    ("lorenz_rk45", L.RK45, lambda: W.lorenz_batched(3000), (3000, 700)),
    ("lorenz_rk23", L.RK23, lambda: W.lorenz_plain(640), (640, 129)),
    ("linear_mfma", L.RK45, lambda: W.linear_shared(40), (40, 17)),
]


@pytest.mark.parametrize("name,method,make,sizes", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_allocating_entry(gpu, name, method, make, sizes):
    import torch
    w = make()
    if name == "lorenz_rk23":
        w["t_end"] = np.full(w["y0"].shape[1], 1.0)
    bufs = [_buffers(torch, w, n) for n in sizes]
    # two allocation-free calls in a row, the second smaller, then the
    # allocating entry point on the same inputs
    got = [_solve(torch, "bg_rk_solve_device_noalloc", method, w, b, n)
           for b, n in zip(bufs, sizes)]
    ref = [_solve(torch, "bg_rk_solve_device", method, w, b, n) for b, n in zip(bufs, sizes)]
    for g, r, n in zip(got, ref, sizes):
        assert np.all(r["st"] == 0)
        for k in ("t", "y", "st", "cnt"):
            assert np.array_equal(g[k], r[k]), (name, n, k)


def test_noalloc_rejects_like_allocating_entry(gpu):
    d = L.EnsembleDesc(n=4, dim=3, n_params=3)  # no y0
    for entry in ("bg_rk_solve_device", "bg_rk_solve_device_noalloc"):
        assert getattr(L.lib(), entry)(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d), None,
                                       None) == L.ERR_INVALID_ARGUMENT
        assert getattr(L.lib(), entry)(L.BDF6, L.RHS["lorenz"], L.INTENDED, C.byref(d), None,
                                       None) == L.ERR_INVALID_ARGUMENT
