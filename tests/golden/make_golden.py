"""Generate the golden fixtures under tests/golden/ (run in the build container).

Independent high-accuracy solutions from scipy (not from the reference, which
is a Rust crate that cannot be built here, and not from this repository's
oracle), used to pin the oracle's "intended" solvers:

  lorenz_golden.npz : 32 Lorenz-63 instances of the plain-ensemble distribution,
                      DOP853 rtol=atol=1e-13, states at t = 0.5, 1, 2, 5
  lv_golden.npz     : 32 Lotka-Volterra instances of the adjoint workload,
                      DOP853 rtol=atol=1e-13, states at t_k = 0.75 k, k=1..20
  bruss_golden.npz  : 8 Brusselator (16 cells) instances, Radau rtol=atol=1e-12,
                      states at t = 1, 10
  robertson_golden.npz : 16 Robertson instances of the stiff-kinetics workload,
                      Radau rtol=1e-12, atol=1e-14, states at t = 10

    python tests/golden/make_golden.py [robertson]
"""
from __future__ import annotations

import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from bacon_amd import workloads as W  # noqa: E402


def lorenz_f(t, y, s, r, b):
    return [s * (y[1] - y[0]), y[0] * (r - y[2]) - y[1], y[0] * y[1] - b * y[2]]


def lv_f(t, y, a, b, g, d):
    return [a * y[0] - b * y[0] * y[1], d * y[0] * y[1] - g * y[1]]


def bruss_f(t, y, A, B, alpha):
    n = y.size // 2
    u, v = y[:n], y[n:]
    c = alpha * n * n
    lu = np.roll(u, 1) - 2 * u + np.roll(u, -1)
    lv = np.roll(v, 1) - 2 * v + np.roll(v, -1)
    uuv = u * u * v
    return np.concatenate([A + uuv - (B + 1) * u + c * lu, B * u - uuv + c * lv])


def robertson_f(t, y, k1, k2, k3):
    return [-k1 * y[0] + k3 * y[1] * y[2], k1 * y[0] - k3 * y[1] * y[2] - k2 * y[1] ** 2,
            k2 * y[1] ** 2]


def robertson():
    w = W.robertson(16)
    out = np.zeros((3, 16))
    for i in range(16):
        s = solve_ivp(robertson_f, (0.0, 10.0), w["y0"][:, i], method="Radau", rtol=1e-12,
                      atol=1e-14, args=tuple(w["params"][:, i]))
        out[:, i] = s.y[:, -1]
    np.savez(os.path.join(HERE, "robertson_golden.npz"), params=w["params"], y=out)


def main():
    robertson()
    w = W.lorenz_plain(32)
    times = np.array([0.5, 1.0, 2.0, 5.0])
    out = np.zeros((32, times.size, 3))
    for i in range(32):
        s = solve_ivp(lorenz_f, (0.0, times[-1]), w["y0"][:, i], method="DOP853",
                      rtol=1e-13, atol=1e-13, t_eval=times, args=tuple(w["params"][:, i]))
        out[i] = s.y.T
    np.savez(os.path.join(HERE, "lorenz_golden.npz"), y0=w["y0"], params=w["params"],
             times=times, y=out)

    w = W.lotka_volterra(32)
    obs = w["obs"]
    out = np.zeros((32, obs.size, 2))
    for i in range(32):
        s = solve_ivp(lv_f, (0.0, obs[-1]), w["y0"][:, i], method="DOP853",
                      rtol=1e-13, atol=1e-13, t_eval=obs, args=tuple(w["params"][:, i]))
        out[i] = s.y.T
    np.savez(os.path.join(HERE, "lv_golden.npz"), y0=w["y0"], params=w["params"],
             times=obs, y=out)

    w = W.brusselator(8)
    times = np.array([1.0, 10.0])
    out = np.zeros((8, times.size, 32))
    for i in range(8):
        s = solve_ivp(bruss_f, (0.0, 10.0), w["y0"][:, i], method="Radau", rtol=1e-12,
                      atol=1e-12, t_eval=times, args=tuple(w["params"][:, i]))
        out[i] = s.y.T
    np.savez(os.path.join(HERE, "bruss_golden.npz"), y0=w["y0"], params=w["params"],
             times=times, y=out)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    if sys.argv[1:] == ["robertson"]:
        robertson()
    else:
        main()
