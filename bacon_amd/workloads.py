"""Seeded synthetic ensembles for the five ensemble workloads.

Every random draw comes from a counter-based SplitMix64 stream (Steele, Lea &
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014):
seed = 0xBAC0_0000 + workload index, stream = instance index, draw k of
instance i is mix64(seed_state(i) + (k+1) * GAMMA).  Uniforms take the top 53
bits; log-uniform values are exp(U(ln a, ln b)); normals use Box-Muller.

Arrays are structure-of-arrays (component-major, shape [dim, n]) — the HBM
layout the kernels read with one coalesced load per component.
"""
from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
SEED_BASE = 0xBAC00000


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(27)
        z *= np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return z


class Stream:
    """Per-instance SplitMix64 streams; draw() returns one value per instance."""

    def __init__(self, workload_index: int, n: int, first: int = 0):
        seed = np.uint64(SEED_BASE + workload_index)
        idx = np.arange(first, first + n, dtype=np.uint64)
        with np.errstate(over="ignore"):
            self.state = _mix64(seed * GAMMA + _mix64(idx + np.uint64(1)))
        self.k = 0

    def bits(self) -> np.ndarray:
        self.k += 1
        with np.errstate(over="ignore"):
            return _mix64(self.state + np.uint64(self.k) * GAMMA)

    def uniform(self, a=0.0, b=1.0) -> np.ndarray:
        u = (self.bits() >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return a + (b - a) * u

    def log_uniform(self, a, b) -> np.ndarray:
        return np.exp(self.uniform(np.log(a), np.log(b)))

    def normal(self) -> np.ndarray:
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def lorenz_plain(n: int = 1024, first: int = 0):
    """Plain RK45 Lorenz-63 ensemble: sigma=10, beta=8/3, rho~U[20,40],
    y0~U[-15,15]xU[-15,15]xU[5,40], t in [0,10], tol 1e-6, dt_min 1e-6,
    dt_max 0.1 (workload index 0)."""
    s = Stream(0, n, first)
    rho = s.uniform(20.0, 40.0)
    y0 = np.stack([s.uniform(-15, 15), s.uniform(-15, 15), s.uniform(5, 40)])
    params = np.stack([np.full(n, 10.0), rho, np.full(n, 8.0 / 3.0)])
    return dict(rhs="lorenz", dim=3, y0=y0, params=params, t0=0.0,
                t_end=np.full(n, 10.0), tol=np.full(n, 1e-6), dt_min=1e-6,
                dt_max=0.1)


def lorenz_batched(n: int = 1 << 20, first: int = 0):
    """Batched RK45 Lorenz-63 ensemble: same parameter and initial-condition
    distributions, t_end~U[1,20], tol log-uniform in [1e-10,1e-4], dt_max 0.1,
    dt_min 1e-12 (not fixed by the workload; chosen and stated here)
    (workload index 1)."""
    s = Stream(1, n, first)
    rho = s.uniform(20.0, 40.0)
    y0 = np.stack([s.uniform(-15, 15), s.uniform(-15, 15), s.uniform(5, 40)])
    t_end = s.uniform(1.0, 20.0)
    tol = s.log_uniform(1e-10, 1e-4)
    params = np.stack([np.full(n, 10.0), rho, np.full(n, 8.0 / 3.0)])
    return dict(rhs="lorenz", dim=3, y0=y0, params=params, t0=0.0, t_end=t_end,
                tol=tol, dt_min=1e-12, dt_max=0.1)


def lotka_volterra(n: int = 262144, first: int = 0, n_obs: int = 20):
    """Lotka-Volterra adjoint ensemble: alpha, beta, gamma, delta log-uniform
    in [0.5,2], y0~U[0.5,2]^2, t in [0,15], tol 1e-8, observations at
    t_k = 0.75 k (k=1..20).  Targets: the parameters scaled by a seeded factor
    in [0.9,1.1] each (the targets are the solution at those perturbed
    parameters, produced by the caller) (workload index 4)."""
    s = Stream(4, n, first)
    params = np.stack([s.log_uniform(0.5, 2.0) for _ in range(4)])
    y0 = np.stack([s.uniform(0.5, 2.0), s.uniform(0.5, 2.0)])
    perturb = np.stack([s.uniform(0.9, 1.1) for _ in range(4)])
    obs = 0.75 * np.arange(1, n_obs + 1)
    return dict(rhs="lotka_volterra", dim=2, y0=y0, params=params, t0=0.0,
                t_end=np.full(n, 15.0), tol=np.full(n, 1e-8), dt_min=1e-12,
                dt_max=0.1, obs=obs, target_params=params * perturb)


def brusselator(n: int = 131072, first: int = 0, cells: int = 16):
    """Stiff 1-D Brusselator: 16 cells x (u, v), periodic, dx = 1/16, A=1,
    B log-uniform in [2,4], alpha=0.02, u0 = A + 0.1 N(0,1),
    v0 = B/A + 0.1 N(0,1) per cell, t in [0,10], tol 1e-7 (workload index 3)."""
    s = Stream(3, n, first)
    A = 1.0
    B = s.log_uniform(2.0, 4.0)
    u0 = np.stack([A + 0.1 * s.normal() for _ in range(cells)])
    v0 = np.stack([B / A + 0.1 * s.normal() for _ in range(cells)])
    y0 = np.concatenate([u0, v0])
    params = np.stack([np.full(n, A), B, np.full(n, 0.02)])
    return dict(rhs="brusselator", dim=2 * cells, y0=y0, params=params, t0=0.0,
                t_end=np.full(n, 10.0), tol=np.full(n, 1e-7), dt_min=1e-12,
                dt_max=0.1)


def robertson(n: int = 4096, first: int = 0):
    """Robertson's stiff chemical kinetics: k1 = 0.04 x logU[0.5, 2],
    k2 = 3e7 x logU[0.5, 2], k3 = 1e4 x logU[0.5, 2], y0 = (1, 0, 0),
    t in [0, 10], tol 1e-6, dt in [1e-10, 5e-4] (the BDF bootstrap is
    explicit RK4, and BDF doubles dt up to dt_max after every accurate step:
    dt_max 5e-4 keeps that bootstrap stable across the rate constants)
    (workload index 5)."""
    s = Stream(5, n, first)
    params = np.stack([0.04 * s.log_uniform(0.5, 2.0), 3e7 * s.log_uniform(0.5, 2.0),
                       1e4 * s.log_uniform(0.5, 2.0)])
    y0 = np.stack([np.ones(n), np.zeros(n), np.zeros(n)])
    return dict(rhs="robertson", dim=3, y0=y0, params=params, t0=0.0,
                t_end=np.full(n, 10.0), tol=np.full(n, 1e-6), dt_min=1e-10,
                dt_max=5e-4)


def two_level(n: int = 4096, first: int = 0):
    """Complex two-level systems c' = (-i H - gamma) c, H = [[hd, ho], [ho,
    -hd]]: hd, ho ~ U[-2, 2], gamma ~ logU[0.01, 0.2]; c0 a random unit
    vector of C^2 stored as (re, im) pairs (dim 4); t in [0, 10], tol 1e-8,
    dt in [1e-12, 0.1] (workload index 6)."""
    s = Stream(6, n, first)
    params = np.stack([s.uniform(-2.0, 2.0), s.uniform(-2.0, 2.0), s.log_uniform(0.01, 0.2)])
    y0 = np.stack([s.normal() for _ in range(4)])
    y0 /= np.sqrt((y0 ** 2).sum(0))
    return dict(rhs="two_level", dim=4, y0=y0, params=params, t0=0.0,
                t_end=np.full(n, 10.0), tol=np.full(n, 1e-8), dt_min=1e-12, dt_max=0.1)


def two_level_exact(params, y0, t):
    """c(t) = exp(-gamma t) (cos(w t) I - i sin(w t) H / w) c0, w = |(hd, ho)|,
    per instance (params [3, n], y0 [4, n] as (re, im) pairs); returns [4, n]."""
    hd, ho, g = params
    c = np.stack([y0[0] + 1j * y0[1], y0[2] + 1j * y0[3]])
    w = np.hypot(hd, ho)
    co, si = np.cos(w * t), np.sin(w * t) / w
    h0 = hd * c[0] + ho * c[1]
    h1 = ho * c[0] - hd * c[1]
    d = np.exp(-g * t)
    r0 = d * (co * c[0] - 1j * si * h0)
    r1 = d * (co * c[1] - 1j * si * h1)
    return np.stack([r0.real, r0.imag, r1.real, r1.imag])


def linear_shared(n: int = 65536, first: int = 0, N: int = 128):
    """Linear y' = A y + b with one shared A = -(Q D Q^T): Q the sign-fixed
    orthogonal factor of a seeded N(0,1) matrix, D log-uniform in [0.1,10],
    b ~ N(0,1) shared; y0 ~ N(0,1); t in [0,5], tol 1e-9 (workload index 2).
    Also returns the eigen-data for the closed-form solution and the shared
    block [A row-major; b] the linear right-hand side reads."""
    g = Stream(2, N * N + 2 * N, 0)  # shared draws use stream indices 0..
    flat = g.normal()
    M = flat[: N * N].reshape(N, N)
    Q, R = np.linalg.qr(M)
    Q = Q * np.sign(np.diag(R))[None, :]
    Dv = np.exp(np.log(0.1) + (np.log(10.0) - np.log(0.1)) * g.uniform()[N * N: N * N + N])
    bvec = g.normal()[N * N + N: N * N + 2 * N]
    A = -(Q * Dv[None, :]) @ Q.T
    s = Stream(2, n, first + (1 << 40))
    y0 = np.stack([s.normal() for _ in range(N)])
    return dict(rhs="linear", dim=N, y0=y0, params=np.zeros((0, n)), t0=0.0,
                t_end=np.full(n, 5.0), tol=np.full(n, 1e-9), dt_min=1e-12,
                dt_max=0.1, A=A, b=bvec, Q=Q, D=Dv,
                shared=np.concatenate([A.reshape(-1), bvec]))


def cost_order(t_end: np.ndarray, tol: np.ndarray, t0: float = 0.0) -> np.ndarray:
    """Longest-first (LPT) schedule: instances sorted by predicted cost
    (t_end - t0) * tol^(-1/4), descending."""
    cost = (t_end - t0) * np.power(tol, -0.25)
    return np.argsort(-cost, kind="stable").astype(np.int32)
