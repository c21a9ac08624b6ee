"""Benchmark: batched RK45 Lorenz-63 ensemble on one MI355X (+ adjoint, BDF6).

Headline step: one batched solve of 1,048,576 Lorenz-63 instances
(sigma=10, beta=8/3, rho~U[20,40], y0~U[-15,15]^2 x U[5,40], t_end~U[1,20],
tol log-uniform in [1e-10,1e-4], dt_max 0.1, dt_min 1e-12), FP64, one
persistent launch, inputs resident in HBM before the timed region.

The JSON line also carries:
  * roofline of the dominant kernel (HIP events on its launch stream) and the
    HBM bytes measured by rocprofv3 (see profiles/);
  * the plain per-instance GPU path on a subset of the same workload;
  * the same ensemble at a uniform tol 1e-8 (the metric's tolerance), with
    its own plain-path comparison;
  * adjoint gradients/sec: discrete adjoint of 262,144 Lotka-Volterra
    instances (20 observations, checkpoint every 16 accepted steps);
  * batched BDF6 on 131,072 stiff Brusselator instances, Jacobian inverse in
    FP64 and in FP32;
  * batched RK45 on 65,536 linear systems (N = 128) sharing one matrix, the
    stage right-hand sides as FP64 MFMA GEMMs;
  * the other methods and outputs on one shared Lorenz workload: RK23,
    Adams5, path output, the solve_ivp chain, and BDF2;
  * a captured HIP graph's replay against the eager entry point, on short
    BDF6 and RK45 solves;
  * the CPU oracle timed on the host cores over a bounded sample.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK45 ensemble solves/sec at tol 1e-8 vs plain GPU path; adjoint grads/sec"
PLAIN_SAMPLE = 1024  # instances timed on the plain per-instance path
PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix peaks are equal)
PEAK_HBM_GBS = 8000.0
# algorithmic FLOPs of one RK45 step attempt (fma = 2):
# 2D(nnz(A)+nnz(e)+nnz(b)) + 6(F_rhs + D) + (2D + 2) + F_ctrl
FLOPS_ATTEMPT_LORENZ = 2 * 3 * (15 + 5 + 4) + 6 * (8 + 3) + 8 + 6   # 224
FLOPS_ATTEMPT_LV = 2 * 2 * (15 + 5 + 4) + 6 * (7 + 2) + 6 + 6       # 162
# Bogacki-Shampine 2(3): nnz(A) = 5, nnz(e) = 4, nnz(b) = 3, 4 stages
FLOPS_ATTEMPT_LORENZ_RK23 = 2 * 3 * (5 + 4 + 3) + 4 * (8 + 3) + 8 + 6  # 130
# replay of one accepted LV step with its recorded size (backward pass): the
# stages and the update, 2D(nnz(A) + nnz(b)) + 6(F_rhs + D)
FLOPS_REPLAY_STEP_LV = 2 * 2 * (15 + 4) + 6 * (7 + 2)                # 130
# reverse of one accepted LV step: stage recompute (6 RHS, 15 nnz(A) combos,
# 6 dt scalings) + 6 VJPs + stage-adjoint combos + b-weighting + sums
FLOPS_VJP_STEP_LV = (6 * 7 + 2 * 2 * 15 + 6 * 2) + 6 * 20 + 2 * 2 * 15 + 2 * 2 * 4 + 6 * 2 * 2
# BDF6 Brusselator (N = 32): full RHS 304 flops; residual = RHS + 15 N
BRUSS_N = 32
F_RHS_BRUSS = 16 * 19
F_RESID = F_RHS_BRUSS + 15 * BRUSS_N
# Per Jacobian inverse as executed: the residual at the guess, the
# finite-difference Jacobian from 4 column colours x 2 signs (each pass one
# residual evaluation of every row), the Gauss-Jordan inverse and the first
# J^-1 F product.
BRUSS_COLOURS = 4
FLOPS_SECANT_SETUP = ((1 + 2 * BRUSS_COLOURS) * F_RESID
                      + 2 * BRUSS_N ** 3 + 2 * BRUSS_N ** 2)
# The same with the reference's finite differences, one column at a time
# (bdf.rs:390-411: 2N residual evaluations): the method's own work.
FLOPS_SECANT_SETUP_REFERENCE_FD = ((1 + 2 * BRUSS_N) * F_RESID
                                   + 2 * BRUSS_N ** 3 + 2 * BRUSS_N ** 2)
# Per Broyden iteration, the reference's way (bdf.rs:449-470): the residual,
# J^-1 (F - F_l), s^T J^-1, the rank-1 update and J^-1 F, plus vector work.
FLOPS_BROYDEN_ITER_REFERENCE = F_RESID + 5 * 2 * BRUSS_N ** 2 + 6 * BRUSS_N
# As executed: J^-1 F after the update from the rank-1 identity (a dot
# product u . F and one fma per component) instead of a fifth product.
FLOPS_BROYDEN_ITER = F_RESID + 4 * 2 * BRUSS_N ** 2 + 10 * BRUSS_N
# linear y' = A y + b, N = 128, one shared A: RHS = matvec (2 N^2) + b (N)
LIN_N = 128
FLOPS_ATTEMPT_LINEAR = (2 * LIN_N * (15 + 5 + 4) + 6 * (2 * LIN_N * LIN_N + LIN_N + LIN_N)
                        + 2 * LIN_N + 2 + 6)
# HBM bytes per Lorenz instance: loads y0(3) rho sigma beta t_end tol, order
# index; stores t, y(3), status, counts
BYTES_PER_INSTANCE = 8 * 8 + 4 + 8 * 4 + 4 + 8
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE and WRITE_SIZE (each
# scaled by a 1 GiB copy-kernel calibration); see
# profiles/rk_ensemble_lorenz_2p20.md
MEASURED_TRAFFIC_BYTES = 1.065e9 + 0.237e9


_BASELINE_ORACLE = None


def baseline_oracle():
    """The CPU oracle in its baseline build (-O3 -march=native, compiled once
    per run on this host)."""
    global _BASELINE_ORACLE
    if _BASELINE_ORACLE is None:
        import oracle
        oracle.use_native_build()
        _BASELINE_ORACLE = oracle
    return _BASELINE_ORACLE


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Ctx:
    def __init__(self, torch, L, dev):
        self.torch, self.L, self.dev = torch, L, dev
        self.lib = L.lib()
        self.stream = torch.cuda.current_stream()
        self.sh = C.c_void_p(self.stream.cuda_stream)

    def T(self, a, dt=None):
        torch = self.torch
        return torch.as_tensor(np.array(a), dtype=dt or torch.float64, device=self.dev)

    def zeros(self, shape, dt=None):
        return self.torch.zeros(shape, dtype=dt or self.torch.float64, device=self.dev)


def rk_ensemble(cx, w, order=True):
    """Device buffers and descriptor for a Lorenz workload dict."""
    from bacon_amd import workloads as W
    torch = cx.torch
    n = w["y0"].shape[1]
    b = dict(y0=cx.T(w["y0"]), params=cx.T(w["params"]), tend=cx.T(w["t_end"]),
             tol=cx.T(w["tol"]), t_out=cx.zeros(n), y_out=cx.zeros((3, n)),
             status=cx.zeros(n, torch.int32), counts=cx.zeros((n, 2), torch.int32))
    if order:
        b["order"] = cx.T(W.cost_order(w["t_end"], w["tol"]), torch.int32)
    d = cx.L.EnsembleDesc(
        n=n, dim=3, n_params=3, y0=b["y0"].data_ptr(), params=b["params"].data_ptr(),
        t0_s=0.0, t_end=b["tend"].data_ptr(), tol=b["tol"].data_ptr(), dt_min_s=w["dt_min"],
        dt_max_s=w["dt_max"], order=b["order"].data_ptr() if order else None,
        t_out=b["t_out"].data_ptr(), y_out=b["y_out"].data_ptr(),
        status=b["status"].data_ptr(), counts=b["counts"].data_ptr())
    return b, d


def graph_replay(cx):
    """One captured graph replay against the eager entry point, on short
    solves where the launch is a visible share of the time: 256 Brusselator
    instances under BDF6 to t = 0.05 (one wave each), and 4096 Lorenz-63
    instances under RK45 to t = 0.1.  The eager call takes its workspace from
    the stream-ordered allocator and launches a setup kernel and the solve
    kernel; the graph owns its workspace and replays both kernels.  Each
    side: 50 back-to-back launches on one stream, after 3 warm-up launches;
    the replay's outputs are checked bit for bit against the eager solve."""
    from bacon_amd import workloads as W
    L, torch = cx.L, cx.torch
    reps = 50
    out = {}

    def bufs(w, n, dim):
        return dict(y0=cx.T(w["y0"]), params=cx.T(w["params"]), t_out=cx.zeros(n),
                    y_out=cx.zeros((dim, n)), status=cx.zeros(n, torch.int32),
                    counts=cx.zeros((n, 2), torch.int32))

    def desc(b, n, dim, t_end, tol, dt_min):
        return L.EnsembleDesc(
            n=n, dim=dim, n_params=3, y0=b["y0"].data_ptr(), params=b["params"].data_ptr(),
            t0_s=0.0, t_end_s=t_end, tol_s=tol, dt_min_s=dt_min, dt_max_s=0.1,
            t_out=b["t_out"].data_ptr(), y_out=b["y_out"].data_ptr(),
            status=b["status"].data_ptr(), counts=b["counts"].data_ptr())

    cases = []
    nb = 256
    bb = bufs(W.brusselator(nb), nb, 32)
    db = desc(bb, nb, 32, 0.05, 1e-7, 1e-12)
    cases.append(("bdf6_brusselator_256_t0.05", bb, db,
                  lambda: cx.lib.bg_bdf_solve_device(L.BDF6, L.RHS["brusselator"], L.INTENDED,
                                                     64, C.byref(db), cx.sh),
                  lambda g: cx.lib.bg_graph_capture_bdf(L.BDF6, L.RHS["brusselator"],
                                                        L.INTENDED, 64, C.byref(db),
                                                        C.byref(g))))
    nr = 4096
    br = bufs(W.lorenz_plain(nr), nr, 3)
    dr = desc(br, nr, 3, 0.1, 1e-6, 1e-6)
    lc = L.Launch(block=256, blocks_per_cu=2)
    cases.append(("rk45_lorenz_4096_t0.1", br, dr,
                  lambda: cx.lib.bg_rk_solve_device(L.RK45, L.RHS["lorenz"], L.INTENDED,
                                                    C.byref(dr), C.byref(lc), cx.sh),
                  lambda g: cx.lib.bg_graph_capture_rk(L.RK45, L.RHS["lorenz"], L.INTENDED,
                                                       C.byref(dr), C.byref(lc), C.byref(g))))
    for name, b, d, eager, capture in cases:
        def run_eager():
            rc = eager()
            if rc != 0:
                raise RuntimeError(f"{name}: eager solve returned {rc}")
        wall_e, kern_e = time_launches(cx, run_eager, reps, 3)
        ref = b["y_out"].cpu().numpy().copy()
        g = C.c_void_p()
        rc = capture(g)
        if rc != 0:
            raise RuntimeError(f"{name}: capture returned {rc}")
        try:
            def run_graph():
                r = cx.lib.bg_graph_launch(g, cx.sh)
                if r != 0:
                    raise RuntimeError(f"{name}: replay returned {r}")
            b["y_out"].zero_()
            wall_g, kern_g = time_launches(cx, run_graph, reps, 3)
            same = bool(np.array_equal(b["y_out"].cpu().numpy(), ref))
        finally:
            cx.lib.bg_graph_destroy(g)
        out[name] = {"eager_us_per_solve": wall_e * 1e6, "graph_us_per_solve": wall_g * 1e6,
                     "eager_event_us": kern_e * 1e6, "graph_event_us": kern_g * 1e6,
                     "replay_identical_to_eager": same}
    out["method"] = (f"{reps} back-to-back launches per side on one stream; us_per_solve = host "
                     "wall time / launches; event_us = HIP events around each launch")
    return out


def autotune(cx, d, args):
    """Tune the plan for descriptor d (d.order must be unset)."""
    L = cx.L
    tn = L.Tuning()
    t = time.perf_counter()
    rc = cx.lib.bg_rk_autotune_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d),
                                      C.byref(tn), cx.sh)
    if rc != 0:
        raise RuntimeError(f"bg_rk_autotune_device returned {rc}")
    tn.pad = 0
    log(f"[bench] autotune {time.perf_counter() - t:.2f} s: {tn.as_dict()}")
    if args.block:
        tn.launch.block = args.block
    if args.blocks_per_cu:
        tn.launch.blocks_per_cu = args.blocks_per_cu
    if args.refill_min:
        tn.launch.refill_min = args.refill_min
    return tn


def time_launches(cx, fn, steps, warmup):
    torch = cx.torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(cx.stream)
        fn()
        e.record(cx.stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean([s.elapsed_time(e) for s, e in ev])) / 1000.0


def cpu_model():
    """The host CPU: /proc/cpuinfo's model name, family, model and stepping
    (a virtualised host may report only the vendor's family name), and the
    CPU flags that decide the -march=native build's vector width."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if info:
                        break
                    continue
                k, v = (s.strip() for s in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "flags") and k not in info:
                    info[k] = v
    except OSError:
        return "unknown"
    flags = set(info.get("flags", "").split())
    simd = [f for f in ("avx2", "fma", "avx512f", "avx512vl", "avx512dq") if f in flags]
    return (f"{info.get('model name', 'unknown')} (family {info.get('cpu family', '?')}, "
            f"model {info.get('model', '?')}, stepping {info.get('stepping', '?')}; "
            f"{' '.join(simd)})")


def plain_path(cx, w, m):
    """The plain per-instance path on the first m instances of workload w."""
    n = w["y0"].shape[1]
    sub = {k: (v[..., :m] if isinstance(v, np.ndarray) and v.ndim and v.shape[-1] == n else v)
           for k, v in w.items()}
    b, d = rk_ensemble(cx, sub, order=False)
    cx.torch.cuda.synchronize()
    t = time.perf_counter()
    rc = cx.lib.bg_rk_solve_plain_device(cx.L.RK45, cx.L.RHS["lorenz"], cx.L.INTENDED,
                                         C.byref(d), cx.sh)
    cx.torch.cuda.synchronize()
    t = time.perf_counter() - t
    if rc != 0:
        raise RuntimeError(f"plain path returned {rc}")
    return m / t, b


def tight_error(cx, w, y_out, m=256, rhs="lorenz"):
    """Final-state error of the first m instances against the same instances
    solved at tol 1e-13 (FP64, this library's RK45, dt_min 1e-15).  y_out:
    the run's [dim][n] final states on the device."""
    L = cx.L
    n = w["y0"].shape[1]
    sub = {k: (v[..., :m] if isinstance(v, np.ndarray) and v.ndim and v.shape[-1] == n else v)
           for k, v in w.items()}
    dim = w["dim"]
    bt = dict(y0=cx.T(sub["y0"]), params=cx.T(sub["params"]),
              tend=cx.T(np.broadcast_to(np.asarray(sub["t_end"], dtype=np.float64), (m,))),
              t=cx.zeros(m), y=cx.zeros((dim, m)), st=cx.zeros(m, cx.torch.int32),
              cnt=cx.zeros((m, 2), cx.torch.int32))
    dt = L.EnsembleDesc(
        n=m, dim=dim, n_params=bt["params"].shape[0], y0=bt["y0"].data_ptr(),
        params=bt["params"].data_ptr(), t0_s=0.0, t_end=bt["tend"].data_ptr(), tol_s=1e-13,
        dt_min_s=1e-15, dt_max_s=w["dt_max"], t_out=bt["t"].data_ptr(),
        y_out=bt["y"].data_ptr(), status=bt["st"].data_ptr(), counts=bt["cnt"].data_ptr())
    rc = cx.lib.bg_rk_solve_device(L.RK45, L.RHS[rhs], L.INTENDED, C.byref(dt), None, cx.sh)
    if rc != 0:
        raise RuntimeError(f"tight solve returned {rc}")
    yt = bt["y"].cpu().numpy()
    yr = y_out[:, :m].cpu().numpy()
    err = np.abs(yr - yt).max(axis=0)
    rel = err / np.maximum(1.0, np.abs(yt).max(axis=0))
    return {"reference": "the same instances at tol 1e-13 (FP64 RK45, this library)",
            "sample": f"first {m} instances",
            "tight_solve_failed": int(np.count_nonzero(bt["st"].cpu().numpy())),
            "max_abs": float(err.max()), "median_abs": float(np.median(err)),
            "max_rel": float(rel.max()), "median_rel": float(np.median(rel))}


def method_family(cx, args):
    """The methods and outputs beyond the headline on one shared workload:
    262,144 plain Lorenz-63 instances (t in [0, 10], tol 1e-6, dt in
    [1e-6, 0.1]).  RK45 as the reference point; RK23; Adams5; RK45 with every
    accepted step recorded in a 64-entry ring per instance; the solve_ivp
    chain (Adams5, then RK45, then BDF6) with a 5000-step cap that sends the
    longer Adams5 instances on to RK45; BDF2 on 16,384 Brusselator instances
    to t = 2 at tol 1e-5 (second order needs ~15,000 steps there at 1e-7).
    RK23 runs to t = 1: a third-order pair at tol 1e-6 takes ~10^5 attempts
    per unit time on this system.  Inputs resident in HBM except for solve_ivp, which runs
    through the host-array builder."""
    from bacon_amd import ivp
    from bacon_amd import workloads as W
    torch, L = cx.torch, cx.L
    nf = 262144
    log(f"[bench] method family, {nf} plain Lorenz instances")
    wf = W.lorenz_plain(nf)
    bf, df = rk_ensemble(cx, wf, order=False)
    out = {"workload": "262,144 Lorenz-63 instances, sigma 10, beta 8/3, rho~U[20,40], "
                       "t in [0,10], tol 1e-6, dt_min 1e-6, dt_max 0.1"}

    def entry(wall, kern, flops_per_attempt=None):
        c = bf["counts"].cpu().numpy().astype(np.int64)
        e = {"solves_per_s": nf / wall, "ms_per_solve": wall * 1e3,
             "failed_instances": int(np.count_nonzero(bf["status"].cpu().numpy())),
             "accepted_steps_per_s": int(c[:, 0].sum()) / wall,
             "attempts": int(c.sum())}
        if flops_per_attempt:
            tf = int(c.sum()) * flops_per_attempt / kern / 1e12
            e.update(fp64_tflops=tf, frac_of_fp64_peak=tf / PEAK_FP64_TFLOPS,
                     flops_per_attempt=flops_per_attempt)
        return e

    for name, method, fpa, t_end in (("rk45", L.RK45, FLOPS_ATTEMPT_LORENZ, None),
                                     ("rk23_to_t1", L.RK23, FLOPS_ATTEMPT_LORENZ_RK23, 1.0)):
        dm = L.EnsembleDesc.from_buffer_copy(df)
        if t_end is not None:
            dm.t_end, dm.t_end_s = None, t_end

        def run():
            rc = cx.lib.bg_rk_solve_device(method, L.RHS["lorenz"], L.INTENDED, C.byref(dm),
                                           None, cx.sh)
            if rc != 0:
                raise RuntimeError(rc)
        wall, kern = time_launches(cx, run, 2, 1)
        out[name] = entry(wall, kern, fpa)
        if method == L.RK45:
            out[name]["error_vs_tol_1e-13"] = tight_error(cx, wf, bf["y_out"])

    def run_adams():
        rc = cx.lib.bg_adams_solve_device(L.ADAMS5, L.RHS["lorenz"], L.INTENDED, C.byref(df),
                                          cx.sh)
        if rc != 0:
            raise RuntimeError(rc)
    wall, kern = time_launches(cx, run_adams, 2, 1)
    out["adams5"] = entry(wall, kern)
    out["adams5"]["note"] = "accepted steps count every yield, bootstrap replays included"

    cap = 64
    pb = dict(t=cx.zeros((cap * nf,)), y=cx.zeros((cap * 3 * nf,)),
              len=cx.zeros(nf, torch.int32))
    for key, flags in (("rk45_path_every_step", 0),
                       ("rk45_path_every_step_instance_major", L.PATH_INSTANCE_MAJOR)):
        pd = L.PathDesc(capacity=cap, every=1, flags=flags, path_t=pb["t"].data_ptr(),
                        path_y=pb["y"].data_ptr(), path_len=pb["len"].data_ptr())

        def run_path():
            rc = cx.lib.bg_rk_solve_path_device(L.RK45, L.RHS["lorenz"], L.INTENDED,
                                                C.byref(df), C.byref(pd), cx.sh)
            if rc != 0:
                raise RuntimeError(rc)
        wall, kern = time_launches(cx, run_path, 2, 1)
        e = entry(wall, kern, FLOPS_ATTEMPT_LORENZ)
        recorded = int(pb["len"].cpu().numpy().astype(np.int64).sum())
        e.update(path_entries_per_s=recorded / wall,
                 path_write_gb_per_s=recorded * 32 / kern / 1e9,
                 overhead_vs_rk45=wall / (out["rk45"]["ms_per_solve"] / 1e3) - 1.0,
                 ring_capacity=cap)
        out[key] = e
    del pb

    m = 65536
    sub = {k: (v[..., :m] if isinstance(v, np.ndarray) and v.ndim and v.shape[-1] == nf else v)
           for k, v in wf.items()}
    ens = (ivp.Ensemble(L.ADAMS5, 3, m).with_minimum_dt(sub["dt_min"])
           .with_maximum_dt(sub["dt_max"]).with_tolerance(sub["tol"])
           .with_initial_time(0.0).with_ending_time(sub["t_end"])
           .with_initial_conditions(sub["y0"]).with_derivative("lorenz", sub["params"])
           .with_max_attempts(5000))
    ens.solve_ivp()
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = ens.solve_ivp()
    t = time.perf_counter() - t
    used = r["method_used"]
    out["solve_ivp_chain"] = {
        "instances": m, "solves_per_s": m / t, "ms_per_solve": t * 1e3,
        "step_cap": 5000, "host_arrays": True,
        "finished_by": {"adams5": int((used == L.ADAMS5).sum()),
                        "rk45": int((used == L.RK45).sum()),
                        "bdf6": int((used == L.BDF6).sum())},
        "failed_instances": int(np.count_nonzero(r["status"]))}
    del ens

    nb2 = 16384
    wb = W.brusselator(nb2)
    bb = dict(y0=cx.T(wb["y0"]), params=cx.T(wb["params"]), t_out=cx.zeros(nb2),
              y_out=cx.zeros((32, nb2)), status=cx.zeros(nb2, torch.int32),
              counts=cx.zeros((nb2, 2), torch.int32))
    db = L.EnsembleDesc(
        n=nb2, dim=32, n_params=3, y0=bb["y0"].data_ptr(), params=bb["params"].data_ptr(),
        t0_s=0.0, t_end_s=2.0, tol_s=1e-5, dt_min_s=1e-12, dt_max_s=0.1,
        t_out=bb["t_out"].data_ptr(), y_out=bb["y_out"].data_ptr(),
        status=bb["status"].data_ptr(), counts=bb["counts"].data_ptr())

    def run_b2():
        rc = cx.lib.bg_bdf_solve_device(L.BDF2, L.RHS["brusselator"], L.INTENDED, 64,
                                        C.byref(db), cx.sh)
        if rc != 0:
            raise RuntimeError(rc)
    wall, kern = time_launches(cx, run_b2, 1, 1)
    out["bdf2_brusselator"] = {
        "instances": nb2, "t_end": 2.0, "tol": 1e-5,
        "solves_per_s": nb2 / wall, "ms_per_solve": wall * 1e3,
        "failed_instances": int(np.count_nonzero(bb["status"].cpu().numpy())),
        "accepted_steps": int(bb["counts"].cpu().numpy()[:, 0].sum())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--refill-min", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--extras", default="tol8,adjoint,bdf,linear,family,graph",
                    help="comma list of extra measurements: tol8, adjoint, bdf, linear, "
                         "family, graph")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline ensemble (used for profiling)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the solve, run a plain device copy of this many MiB "
                         "(reference dispatch for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(instances are independent: a multi-GPU run gives each GPU a contiguous "
              "range of the longest-first order, with no exchange between GPUs)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from bacon_amd import _lib as L
    from bacon_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    cx = Ctx(torch, L, dev)
    n = args.n
    log(f"[bench] generating workload n={n}")
    w = W.lorenz_batched(n)
    b, d_full = rk_ensemble(cx, w, order=False)
    # the autotuner fits the schedule model and picks the launch shape once
    # (a plan, outside the timed region); every timed step computes the
    # schedule (cost kernel + radix sort) and solves
    tn = autotune(cx, d_full, args)
    # one step = schedule (bg_rk_schedule_device) + solve (bg_rk_solve_device
    # with that order and the tuned shape), i.e. bg_rk_solve_tuned_device
    # split in two so an event pair brackets the solve kernel alone
    order_buf = cx.zeros(n, torch.int32)
    d_sched = L.EnsembleDesc.from_buffer_copy(d_full)
    d_sched.order = order_buf.data_ptr()
    kern_events = []

    def run_once():
        rc = cx.lib.bg_rk_schedule_device(C.byref(tn), C.byref(d_full),
                                          C.c_void_p(order_buf.data_ptr()), cx.sh)
        if rc != 0:
            raise RuntimeError(f"bg_rk_schedule_device returned {rc}")
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(cx.stream)
        rc = cx.lib.bg_rk_solve_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d_sched),
                                       C.byref(tn.launch), cx.sh)
        e1.record(cx.stream)
        kern_events.append((e0, e1))
        if rc != 0:
            raise RuntimeError(f"bg_rk_solve_device returned {rc}")

    if args.copy_calibration_mib:
        nb = args.copy_calibration_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: {nb} bytes read, {nb} bytes written")
        del src, dst

    log(f"[bench] headline: warmup {args.warmup}, {args.steps} timed steps")
    wall, step_ev = time_launches(cx, run_once, args.steps, args.warmup)
    kern = float(np.mean([a.elapsed_time(b) for a, b in kern_events[-args.steps:]])) / 1000.0
    ms_per_step = wall * 1000.0
    log(f"[bench] ms/step {ms_per_step:.3f} (step events {step_ev * 1e3:.3f}, "
        f"solve kernel {kern * 1e3:.3f})")

    st = b["status"].cpu().numpy()
    cnt = b["counts"].cpu().numpy().astype(np.int64)
    n_acc, n_rej = int(cnt[:, 0].sum()), int(cnt[:, 1].sum())
    attempts = n_acc + n_rej
    per_inst = cnt[:, 0] + cnt[:, 1]
    solves_per_s = n / wall
    achieved = attempts * FLOPS_ATTEMPT_LORENZ / kern / 1e12

    result = {
        "metric": METRIC,
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 ensemble)",
        "config": {
            "workload": "batched RK45 (Fehlberg, intended tableau), 1,048,576 Lorenz-63 "
                        "instances, t_end~U[1,20], tol log-uniform [1e-10,1e-4], "
                        "dt_max 0.1, dt_min 1e-12, one persistent launch",
            "n_instances": n,
            "semantics": "intended",
            "parallelism": "one lane per instance, persistent grid, wave-ballot refill, "
                           "longest-first schedule from the autotuner's fitted cost model "
                           "(computed on device each step), autotuned launch shape",
            "tuning": tn.as_dict(),
        },
        "accepted_steps_per_s": n_acc / wall,
        "attempts_per_launch": attempts,
        "attempts_per_instance": {"mean": float(per_inst.mean()),
                                  "p99": float(np.percentile(per_inst, 99)),
                                  "max": int(per_inst.max())},
        "failed_instances": int(np.count_nonzero(st)),
        "roofline": {
            "bound": "fp64_valu",
            "compute": "fp64 (MI355X fp64 vector peak = fp64 matrix peak = 78.6 TFLOP/s; "
                       "this kernel issues FP64 VALU fma, no MFMA)",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": MEASURED_TRAFFIC_BYTES,
            "traffic_source": "profiles/rk_ensemble_lorenz_2p20.md (rocprofv3 --pmc FETCH_SIZE "
                              "and WRITE_SIZE runs of this workload; a constant here, not "
                              "measured in this run)",
            "traffic_note": "rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (factors from a "
                            "1 GiB copy kernel); above the algorithmic bytes because the "
                            "longest-first schedule gathers each instance's inputs and scatters "
                            "its outputs through order[k], one cache line per 8-byte value: "
                            "~0.26 ms of HBM time in a compute-bound 59 ms kernel",
            "kernel": "rk_ensemble_kernel<RhsLorenz, Tableau<Fehlberg45,false>>",
            "flops_per_attempt": FLOPS_ATTEMPT_LORENZ,
            "avg_kernel_ms": kern * 1e3,
            "step_events_ms": step_ev * 1e3,
            "algorithmic_hbm_bytes": n * BYTES_PER_INSTANCE,
        },
    }

    result["error_vs_tol_1e-13"] = tight_error(cx, w, b["y_out"])
    result["error_vs_tol_1e-13"]["note"] = (
        "Lorenz-63 is chaotic: a local error of the tolerance grows ~e^(0.9 t) by t_end "
        "(up to 20), so late final states differ at O(1) from any other solver's")

    if not args.no_plain:
        m = PLAIN_SAMPLE
        log(f"[bench] plain per-instance path, {m} instances")
        plain_sps, pb = plain_path(cx, w, m)
        same = bool(np.array_equal(pb["y_out"].cpu().numpy(),
                                   b["y_out"][:, :m].cpu().numpy()))
        result["plain_gpu"] = {"solves_per_s": plain_sps, "sample": f"first {m} instances",
                               "identical_to_batched": same}
        result["speedup_vs_plain"] = solves_per_s / plain_sps

    extras = set() if args.no_extras else {x for x in args.extras.split(",") if x}
    if "tol8" in extras:
        # ---- the metric's tolerance: uniform tol 1e-8 ---------------------------
        log("[bench] uniform tol 1e-8 ensemble")
        wu = dict(w)
        wu["tol"] = np.full(n, 1e-8)
        bu, du = rk_ensemble(cx, wu, order=False)
        tnu = autotune(cx, du, args)

        def run_u():
            rc = cx.lib.bg_rk_solve_tuned_device(L.RK45, L.RHS["lorenz"], L.INTENDED,
                                                 C.byref(du), C.byref(tnu), cx.sh)
            if rc != 0:
                raise RuntimeError(rc)
        wall_u, kern_u = time_launches(cx, run_u, args.steps, args.warmup)
        cu = bu["counts"].cpu().numpy().astype(np.int64)
        tol8 = {"solves_per_s": n / wall_u, "ms_per_solve": wall_u * 1e3,
                "accepted_steps_per_s": int(cu[:, 0].sum()) / wall_u,
                "fp64_tflops": int(cu.sum()) * FLOPS_ATTEMPT_LORENZ / kern_u / 1e12,
                "frac_of_fp64_peak": int(cu.sum()) * FLOPS_ATTEMPT_LORENZ / kern_u / 1e12
                                     / PEAK_FP64_TFLOPS,
                "tuning": tnu.as_dict(),
                "error_vs_tol_1e-13": tight_error(cx, wu, bu["y_out"])}
        if not args.no_plain:
            mu = PLAIN_SAMPLE
            plain_u, pbu = plain_path(cx, wu, mu)
            tol8["plain_gpu_solves_per_s"] = plain_u
            tol8["plain_sample"] = f"first {mu} instances"
            tol8["speedup_vs_plain"] = (n / wall_u) / plain_u
            tol8["identical_to_batched"] = bool(np.array_equal(
                pbu["y_out"].cpu().numpy(), bu["y_out"][:, :mu].cpu().numpy()))
        result["uniform_tol_1e-8"] = tol8
        del bu, du

    if "adjoint" in extras:
        # ---- adjoint gradients --------------------------------------------------
        from bacon_amd.adjoint import solve_vjp
        na = 262144
        log(f"[bench] adjoint, {na} Lotka-Volterra instances")
        wl = W.lotka_volterra(na)
        obs = cx.T(wl["obs"])
        zeros = cx.zeros((wl["obs"].size, 2, na))
        tr = solve_vjp("lotka_volterra", cx.T(wl["y0"]), cx.T(wl["target_params"]), obs, zeros,
                       15.0, 1e-8)
        targets = tr["obs_y"].clone()
        del tr
        y0l, pl = cx.T(wl["y0"]), cx.T(wl["params"])
        res = {}

        def run_adj():
            res["r"] = solve_vjp("lotka_volterra", y0l, pl, obs, targets, 15.0, 1e-8)
        wall_a, kern_a = time_launches(cx, run_adj, 3, 1)
        r = res["r"]
        ca = r["counts"].cpu().numpy().astype(np.int64)
        acc_a, att_a = int(ca[:, 0].sum()), int(ca.sum())
        # forward attempts + backward replayed accepted steps + reverse steps
        flops_a = att_a * FLOPS_ATTEMPT_LV + acc_a * (FLOPS_REPLAY_STEP_LV + FLOPS_VJP_STEP_LV)
        result["adjoint_grads_per_s"] = na / wall_a
        result["adjoint"] = {
            "workload": "discrete adjoint of batched RK45, 262,144 Lotka-Volterra instances, "
                        "tol 1e-8, t in [0,15], 20 observations, checkpoint every 16 steps; "
                        "gradients w.r.t. 4 parameters and y0",
            "grads_per_s": na / wall_a, "ms_per_solve": wall_a * 1e3,
            "failed_instances": int(np.count_nonzero(r["status"].cpu().numpy())),
            "accepted_steps": acc_a,
            "fp64_tflops": flops_a / kern_a / 1e12,
        }
        # the tuner's plan (checkpoint interval, resident blocks per CU)
        from bacon_amd.adjoint import autotune_vjp
        atn = autotune_vjp("lotka_volterra", y0l, pl, obs, targets, 15.0, 1e-8)

        def run_adj_tuned():
            res["t"] = solve_vjp("lotka_volterra", y0l, pl, obs, targets, 15.0, 1e-8, plan=atn)
        wall_at, _ = time_launches(cx, run_adj_tuned, 3, 1)
        result["adjoint"]["tuned"] = {
            "plan": atn.as_dict(),
            "default_plan": {"checkpoint_every": 16, "max_checkpoints": 128,
                             "blocks_per_cu": "forward 2, backward 1"},
            "grads_per_s": na / wall_at, "ms_per_solve": wall_at * 1e3,
            "bit_identical_to_default": bool(torch.equal(res["t"]["grad_params"],
                                                         r["grad_params"]))}
        del res["t"]
        # The same gradient solve inside an optimisation loop: each step's
        # parameters are the previous ones perturbed (exp(1e-3 N(0,1)) per
        # entry, a stand-in for an optimiser update), and each step's forward
        # pass runs longest first by the previous step's attempt counts
        # (bg_schedule_by_counts_device, inside the timed step).
        from bacon_amd.adjoint import longest_first
        rng = np.random.default_rng(0xBAC00005)
        loop_params = [cx.T(wl["params"] * np.exp(1e-3 * rng.standard_normal(wl["params"].shape)))
                       for _ in range(4)]
        st = {"k": 0, "order": None}

        def run_loop():
            p = loop_params[st["k"] % len(loop_params)]
            st["k"] += 1
            rl = solve_vjp("lotka_volterra", y0l, p, obs, targets, 15.0, 1e-8, order=st["order"])
            st["order"] = longest_first(rl["counts"])
            res["loop"] = rl
        wall_l, _ = time_launches(cx, run_loop, 3, 1)
        result["adjoint"]["optimisation_loop"] = {
            "grads_per_s": na / wall_l, "ms_per_solve": wall_l * 1e3,
            "failed_instances": int(np.count_nonzero(res["loop"]["status"].cpu().numpy())),
            "schedule": "forward pass longest first by the previous step's attempt counts",
            "params": "perturbed every step by exp(1e-3 N(0,1)) per entry; 4 parameter sets "
                      "cycled, so no step reuses its own counts"}
        del res["loop"], loop_params
        gold = np.load(os.path.join(ROOT, "tests", "golden", "lv_golden.npz"))
        mg = gold["y0"].shape[1]
        if np.array_equal(gold["y0"], wl["y0"][:, :mg]) and np.array_equal(gold["params"],
                                                                         wl["params"][:, :mg]):
            oy = r["obs_y"][:, :, :mg].cpu().numpy()      # [n_obs][2][mg]
            ey = np.abs(oy - gold["y"].transpose(1, 2, 0))
            result["adjoint"]["obs_error_vs_dop853_1e-13"] = {
                "sample": f"first {mg} instances x 20 observations",
                "max_abs": float(ey.max()), "median_abs": float(np.median(ey))}
        del r, res

    if "bdf" in extras:
        # ---- BDF6 Brusselator ---------------------------------------------------
        nb_ = 131072
        log(f"[bench] BDF6, {nb_} Brusselator instances")
        wb = W.brusselator(nb_)
        bb = dict(y0=cx.T(wb["y0"]), params=cx.T(wb["params"]), t_out=cx.zeros(nb_),
                  y_out=cx.zeros((32, nb_)), status=cx.zeros(nb_, torch.int32),
                  counts=cx.zeros((nb_, 2), torch.int32), work=cx.zeros((nb_, 2), torch.int32))
        db = L.EnsembleDesc(
            n=nb_, dim=32, n_params=3, y0=bb["y0"].data_ptr(), params=bb["params"].data_ptr(),
            t0_s=0.0, t_end_s=10.0, tol_s=1e-7, dt_min_s=1e-12, dt_max_s=0.1,
            t_out=bb["t_out"].data_ptr(), y_out=bb["y_out"].data_ptr(),
            status=bb["status"].data_ptr(), counts=bb["counts"].data_ptr(),
            work=bb["work"].data_ptr())
        bdf = {"workload": "batched BDF6 (intended), 131,072 1-D Brusselator instances "
                           "(16 cells x 2 species), B log-uniform [2,4], tol 1e-7, t in [0,10]"}
        y64 = None
        for bits in (64, 32):
            def run_b():
                rc = cx.lib.bg_bdf_solve_device(L.BDF6, L.RHS["brusselator"], L.INTENDED, bits,
                                                C.byref(db), cx.sh)
                if rc != 0:
                    raise RuntimeError(rc)
            wall_b, kern_b = time_launches(cx, run_b, 1, 1)
            wk = bb["work"].cpu().numpy().astype(np.int64)
            flops_b = (int(wk[:, 1].sum()) * FLOPS_SECANT_SETUP
                       + int(wk[:, 0].sum()) * FLOPS_BROYDEN_ITER)
            flops_ref = (int(wk[:, 1].sum()) * FLOPS_SECANT_SETUP_REFERENCE_FD
                         + int(wk[:, 0].sum()) * FLOPS_BROYDEN_ITER_REFERENCE)
            yb = bb["y_out"].cpu().numpy()
            entry = {"solves_per_s": nb_ / wall_b, "ms_per_solve": wall_b * 1e3,
                     "failed_instances": int(np.count_nonzero(bb["status"].cpu().numpy())),
                     "accepted_steps": int(bb["counts"].cpu().numpy()[:, 0].sum()),
                     "broyden_iterations": int(wk[:, 0].sum()),
                     "jacobian_inverses": int(wk[:, 1].sum()),
                     "tflops": flops_b / kern_b / 1e12,
                     "tflops_note": "FP64-equivalent work as executed (coloured finite "
                                    "differences: 9 residual evaluations per Jacobian; "
                                    "four N^2 products per Broyden iteration, J^-1 F "
                                    "from the rank-1 identity); tflops_reference_fd_work "
                                    "counts the reference's 2N-column differences and "
                                    "five products",
                     "tflops_reference_fd_work": flops_ref / kern_b / 1e12}
            if bits == 64:
                y64 = yb
                gold = np.load(os.path.join(ROOT, "tests", "golden", "bruss_golden.npz"))
                mg = gold["y0"].shape[1]
                if np.array_equal(gold["y0"], wb["y0"][:, :mg]):
                    eb = np.abs(yb[:, :mg] - gold["y"][:, 1, :].T)
                    entry["error_vs_radau_1e-12"] = {
                        "sample": f"first {mg} instances at t = 10",
                        "max_abs": float(eb.max()), "median_abs": float(np.median(eb))}
            else:
                entry["max_abs_diff_vs_fp64_lu"] = float(np.abs(yb - y64).max())
            bdf[f"lu_fp{bits}"] = entry
        # the tuner's plan (fitted longest-first schedule + resident-wave cap)
        tn = L.Tuning()
        rc = cx.lib.bg_bdf_autotune_device(L.BDF6, L.RHS["brusselator"], L.INTENDED, 64,
                                           C.byref(db), C.byref(tn), cx.sh)
        if rc != 0:
            raise RuntimeError(rc)

        def run_bt():
            rc = cx.lib.bg_bdf_solve_tuned_device(L.BDF6, L.RHS["brusselator"], L.INTENDED, 64,
                                                  C.byref(db), C.byref(tn), cx.sh)
            if rc != 0:
                raise RuntimeError(rc)
        wall_t, _ = time_launches(cx, run_bt, 1, 1)
        bdf["tuned_lu_fp64"] = {
            "plan": {"waves_per_cu_cap": int(tn.launch.blocks_per_cu) or "occupancy",
                     "schedule": "longest-first by the fitted cost model",
                     "coef": [float(c) for c in tn.coef[:6]]},
            "default_plan": {"waves_per_cu_cap": "occupancy", "schedule": "input order"},
            "tuner_ms": {"pilot_input_order": tn.pilot_ms, "fitted_schedule_occupancy":
                         tn.default_ms, "best": tn.best_ms, "trials": int(tn.n_trials)},
            "ms_per_solve": wall_t * 1e3, "solves_per_s": nb_ / wall_t,
            "bit_identical_to_untuned": bool(np.array_equal(bb["y_out"].cpu().numpy(), y64))}
        result["bdf6_brusselator"] = bdf
        del bb, db

    if "linear" in extras:
        # ---- linear shared-A RK45 (FP64 MFMA tiles) -----------------------------
        nl = 65536
        log(f"[bench] linear shared-A RK45, {nl} instances, N={LIN_N}")
        wl = W.linear_shared(nl)
        bl = dict(y0=cx.T(wl["y0"]), shared=cx.T(wl["shared"]), t_out=cx.zeros(nl),
                  y_out=cx.zeros((LIN_N, nl)), status=cx.zeros(nl, torch.int32),
                  counts=cx.zeros((nl, 2), torch.int32))
        dl = L.EnsembleDesc(
            n=nl, dim=LIN_N, n_params=0, y0=bl["y0"].data_ptr(), shared=bl["shared"].data_ptr(),
            t0_s=0.0, t_end_s=5.0, tol_s=1e-9, dt_min_s=wl["dt_min"], dt_max_s=wl["dt_max"],
            t_out=bl["t_out"].data_ptr(), y_out=bl["y_out"].data_ptr(),
            status=bl["status"].data_ptr(), counts=bl["counts"].data_ptr())

        def run_l():
            rc = cx.lib.bg_rk_solve_device(L.RK45, L.RHS["linear"], L.INTENDED, C.byref(dl),
                                           None, cx.sh)
            if rc != 0:
                raise RuntimeError(rc)
        wall_l, kern_l = time_launches(cx, run_l, 2, 1)
        cl = bl["counts"].cpu().numpy().astype(np.int64)
        yl = bl["y_out"].cpu().numpy()
        xs = np.linalg.solve(wl["A"], -wl["b"])
        ye = xs[:, None] + wl["Q"] @ (np.exp(-wl["D"] * 5.0)[:, None]
                                      * (wl["Q"].T @ (wl["y0"] - xs[:, None])))
        lin = {"workload": "batched RK45 (intended), 65,536 instances of y' = A y + b, "
                           "one shared 128x128 A = -Q diag(D) Q^T, D log-uniform [0.1,10], "
                           "tol 1e-9, t in [0,5]; stage right-hand sides as FP64 MFMA GEMMs",
               "solves_per_s": nl / wall_l, "ms_per_solve": wall_l * 1e3,
               "failed_instances": int(np.count_nonzero(bl["status"].cpu().numpy())),
               "attempts": int(cl.sum()),
               "fp64_tflops": int(cl.sum()) * FLOPS_ATTEMPT_LINEAR / kern_l / 1e12,
               "frac_of_fp64_peak": int(cl.sum()) * FLOPS_ATTEMPT_LINEAR / kern_l / 1e12
                                    / PEAK_FP64_TFLOPS,
               "max_abs_err_vs_closed_form": float(np.abs(yl - ye).max()),
               "mfma_utilisation_rocprof": 0.78,
               "hbm_traffic_bytes_rocprof": 68.4e6 + 68.6e6,
               "rocprof_source": "profiles/linear_mfma_65536.md (rocprofv3 --pmc runs of "
                                 "this workload; constants here, not measured in this run)",
               "algorithmic_hbm_bytes": nl * LIN_N * 8 * 2 + nl * 24}
        if not args.no_cpu_baseline:
            O = baseline_oracle()
            m, threads = 64, min(os.cpu_count() or 1, 16)
            tcl = time.perf_counter()
            O.solve_ensemble(O.RK45, "linear", wl["y0"][:, :m], None, wl["t_end"][:m],
                             wl["tol"][:m], dt_min=wl["dt_min"], dt_max=wl["dt_max"],
                             shared=wl["shared"], semantics=O.INTENDED,
                             profile=O.PROFILE_BACON, threads=threads)
            tcl = time.perf_counter() - tcl
            lin["cpu_oracle"] = {"solves_per_s": m / tcl, "cores": threads,
                                 "sample": f"first {m} instances"}
        result["linear_shared_mfma"] = lin
        del bl, dl

    if "family" in extras:
        result["method_family"] = method_family(cx, args)

    if "graph" in extras:
        log("[bench] graph replay vs eager launch")
        result["graph_replay"] = graph_replay(cx)

    if not args.no_cpu_baseline:
        O = baseline_oracle()
        threads = min(os.cpu_count() or 1, 16)
        log(f"[bench] CPU oracle baseline ({threads} threads, ~{args.cpu_seconds}s)")

        def cpu_runs(nthreads, seconds, chunk):
            """Three runs over one sample: the first solves as many chunks
            of the workload as fit in seconds / 3, the other two the same
            instances again.  Returns (instances, [solves/s per run])."""
            def solve(first):
                sm = W.lorenz_batched(chunk, first=first)
                O.solve_ensemble(O.RK45, "lorenz", sm["y0"], sm["params"], sm["t_end"],
                                 sm["tol"], dt_min=w["dt_min"], dt_max=w["dt_max"],
                                 semantics=O.INTENDED, profile=O.PROFILE_BACON,
                                 threads=nthreads)
            done, t = 0, time.perf_counter()
            while time.perf_counter() - t < seconds / 3 and done < n:
                solve(done)
                done += chunk
            rates = [done / (time.perf_counter() - t)]
            for _ in range(2):
                t = time.perf_counter()
                for first in range(0, done, chunk):
                    solve(first)
                rates.append(done / (time.perf_counter() - t))
            return done, rates

        done, rates = cpu_runs(threads, args.cpu_seconds, 512)
        # bacon itself is single-threaded: the same oracle on one thread, on a
        # shorter sample of the same instances
        done1, rates1 = cpu_runs(1, args.cpu_seconds / 4, 64)
        rate, rate1 = float(np.median(rates)), float(np.median(rates1))
        result["cpu_baseline"] = {
            "value": rate, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"first {done} instances of the same workload, C oracle "
                      f"(bacon arithmetic, gcc -O3 -march=native on this host), OpenMP over "
                      f"{threads} threads; median of 3 "
                      f"runs over that sample, {args.cpu_seconds:.0f} s in all",
            "runs": rates,
            "single_thread": {"value": rate1, "unit": "solves/s", "runs": rates1,
                              "sample": f"first {done1} instances, one thread, median of 3 runs"},
            "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count()},
        }
        result["speedup_vs_cpu"] = solves_per_s / rate
        result["speedup_vs_cpu_single_thread"] = solves_per_s / rate1

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
