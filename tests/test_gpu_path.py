"""Ensemble path output (bg_rk_solve_path_device) against the oracle's path.

The reference's primary output is the iterator's path: every accepted (t, y)
(src/ivp/rk.rs:419, src/ivp.rs:220-238).  The oracle records the same
sequence; the kernel's arithmetic is the ensemble kernel's, so every entry is
compared bit for bit against the oracle's DEVICE profile:
  * every accepted step, on 96 Lorenz-63 instances of the plain workload;
  * with shared stop times (exact landings, the same path with stops added);
  * stop-only recording (dense output at given times);
  * decimation (every k-th step) and ring wrap-around;
  * the instance-major layout holds the same entries.
"""
import numpy as np
import pytest

import oracle as O
from bacon_amd import _lib as L
from bacon_amd import ivp
from bacon_amd import workloads as W

pytestmark = pytest.mark.gpu

N = 96
STOPS = np.array([0.5, 1.25, 2.0, 3.0, 4.5, 6.0, 7.75, 9.5])


@pytest.fixture(scope="module")
def w():
    return W.lorenz_plain(N)


def _oracle(w, i, stops=None):
    return O.solve(O.RK45, "lorenz", w["y0"][:, i], 0.0, 10.0, 1e-6, 1e-6, 0.1,
                   params=w["params"][:, i], semantics=O.INTENDED, profile=O.PROFILE_DEVICE,
                   path_cap=200000, stops=stops)


def test_every_accepted_step_bit_exact(gpu, orc, w):
    g = ivp.solve_path(L.RK45, w, capacity=4096)
    assert np.all(g["status"] == 0)
    for i in range(N):
        r = _oracle(w, i)
        m = r["n_accepted"]
        assert g["path_len"][i] == m == g["n_accepted"][i]
        assert np.array_equal(g["path_t"][:m, i], r["path_t"])
        assert np.array_equal(g["path_y"][:m, :, i], r["path_y"])
        assert np.array_equal(g["y"][:, i], r["y"])


def test_stops_land_exactly(gpu, orc, w):
    g = ivp.solve_path(L.RK45, w, capacity=4096, stops=STOPS)
    for i in range(0, N, 5):
        r = _oracle(w, i, stops=STOPS)
        m = r["n_accepted"]
        assert g["path_len"][i] == m
        assert np.array_equal(g["path_t"][:m, i], r["path_t"])
        assert np.array_equal(g["path_y"][:m, :, i], r["path_y"])
        for k, s in enumerate(STOPS):
            j = np.nonzero(g["path_t"][:m, i] == s)[0]
            assert j.size == 1
            assert np.array_equal(g["path_y"][j[0], :, i], r["stop_y"][k])


def test_stops_only_dense_output(gpu, orc, w):
    g = ivp.solve_path(L.RK45, w, capacity=16, stops=STOPS, stops_only=True)
    assert np.all(g["path_len"] == STOPS.size + 1)
    assert np.all(g["path_t"][: STOPS.size] == STOPS[:, None])
    assert np.all(g["path_t"][STOPS.size] == 10.0)
    for i in range(0, N, 7):
        r = _oracle(w, i, stops=STOPS)
        assert np.array_equal(g["path_y"][: STOPS.size, :, i], r["stop_y"])
        assert np.array_equal(g["path_y"][STOPS.size, :, i], r["y"])


def test_every_kth_and_ring_wrap(gpu, orc, w):
    k = 7
    g = ivp.solve_path(L.RK45, w, capacity=4096, every=k)
    h = ivp.solve_path(L.RK45, w, capacity=16)
    for i in range(0, N, 11):
        r = _oracle(w, i)
        m = r["n_accepted"]
        mk = m // k
        assert g["path_len"][i] == mk
        assert np.array_equal(g["path_t"][:mk, i], r["path_t"][k - 1::k][:mk])
        assert np.array_equal(g["path_y"][:mk, :, i], r["path_y"][k - 1::k][:mk])
        # ring of 16: entry j holds the last recorded step with index = j mod 16
        assert h["path_len"][i] == m
        for j in range(16):
            last = max(q for q in range(m) if q % 16 == j)
            assert h["path_t"][j, i] == r["path_t"][last]
            assert np.array_equal(h["path_y"][j, :, i], r["path_y"][last])


def test_stop_next_to_accepted_step(gpu, orc, w):
    """A stop dt_min/8 after instance 0's 100th accepted time: the landing
    keeps the controller's step size and every instance finishes, bit for bit
    with the oracle's path."""
    k = 100
    stop = _oracle(w, 0)["path_t"][k] + 1e-6 / 8
    stops = np.array([stop, 5.0])
    g = ivp.solve_path(L.RK45, w, capacity=4096, stops=stops)
    assert np.all(g["status"] == 0)
    for i in (0, 1, 2, 50):
        r = _oracle(w, i, stops=stops)
        assert r["status"] == 0
        m = r["n_accepted"]
        assert g["path_len"][i] == m
        assert np.array_equal(g["path_t"][:m, i], r["path_t"])
        assert np.array_equal(g["path_y"][:m, :, i], r["path_y"])
    assert g["path_t"][k + 1, 0] == stop


def test_invalid_arguments(gpu, w):
    import ctypes as C
    d = L.EnsembleDesc(n=4, dim=3, n_params=3)
    p = L.PathDesc(capacity=8)  # no buffers
    assert L.lib().bg_rk_solve_path_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d),
                                           C.byref(p), None) == -1


def test_instance_major_layout_same_entries(gpu, w):
    for kw in (dict(capacity=4096), dict(capacity=16), dict(capacity=16, stops=STOPS,
                                                           stops_only=True)):
        a = ivp.solve_path(L.RK45, w, **kw)
        b = ivp.solve_path(L.RK45, w, instance_major=True, **kw)
        for k in ("path_t", "path_y", "path_len", "y", "t", "status"):
            assert np.array_equal(a[k], b[k]), (kw, k)
