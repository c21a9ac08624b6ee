"""Benchmark: batched RK45 Lorenz-63 ensemble on one MI355X.

A step is one batched solve of the whole ensemble: 1,048,576 Lorenz-63
instances (sigma=10, beta=8/3, rho~U[20,40], y0~U[-15,15]^2 x U[5,40],
t_end~U[1,20], tol log-uniform in [1e-10,1e-4], dt_max 0.1, dt_min 1e-12),
FP64, one persistent launch, inputs resident in HBM before timing starts.

Prints one JSON line: solves/sec (value), accepted steps/sec, the kernel's
FP64 roofline fraction measured with HIP events on the launch stream, the
plain per-instance GPU path on a subset of the same workload, and the CPU
oracle timed on the host cores over a bounded sample.

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
PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix peaks are equal)
PEAK_HBM_GBS = 8000.0
# algorithmic FLOPs of one RK45 step attempt on Lorenz-63 (fma = 2):
# 2D(nnz(A)+nnz(e)+nnz(b)) + 6(F_rhs + D) + (2D + 2) + F_ctrl
#   = 2*3*(15+5+4) + 6*(8+3) + 8 + 6 = 224
FLOPS_PER_ATTEMPT = 224
# HBM bytes per instance: loads y0(3) rho sigma beta t_end tol (+order) and
# stores t, y(3), status, counts
BYTES_PER_INSTANCE = 8 * 8 + 4 + 8 * 4 + 4 + 8
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE and WRITE_SIZE (each
# scaled by a 1 GiB copy-kernel calibration); see
# profiles/rk_ensemble_lorenz_2p20.md
MEASURED_TRAFFIC_BYTES = 1.054e9 + 0.238e9


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the solve, run a plain device copy of this many MiB "
                         "(reference dispatch for HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(instances are independent, see DESIGN.md for how the path shards)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from bacon_amd import _lib as L
    from bacon_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    lib = L.lib()
    n = args.n
    log(f"[bench] generating workload n={n}")
    w = W.lorenz_batched(n)
    order_np = W.cost_order(w["t_end"], w["tol"])

    def T(a, dt=torch.float64):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)

    y0, params = T(w["y0"]), T(w["params"])
    tend, tol = T(w["t_end"]), T(w["tol"])
    order = T(order_np, torch.int32)
    t_out = torch.zeros(n, dtype=torch.float64, device=dev)
    y_out = torch.zeros((3, n), dtype=torch.float64, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    counts = torch.zeros((n, 2), dtype=torch.int32, device=dev)

    def desc(nn, order_ptr=None):
        return L.EnsembleDesc(
            n=nn, dim=3, n_params=3, y0=y0.data_ptr(), params=params.data_ptr(),
            t0_s=0.0, t_end=tend.data_ptr(), tol=tol.data_ptr(), dt_min_s=w["dt_min"],
            dt_max_s=w["dt_max"], order=order_ptr, t_out=t_out.data_ptr(),
            y_out=y_out.data_ptr(), status=status.data_ptr(), counts=counts.data_ptr())

    # a subset descriptor must address the first nn instances with stride nn;
    # the full descriptor uses stride n.
    d_full = desc(n, order.data_ptr())
    launch = L.Launch(block=args.block, blocks_per_cu=args.blocks_per_cu,
                      refill_min=args.refill_min)
    stream = torch.cuda.current_stream()
    sh = C.c_void_p(stream.cuda_stream)

    def run_once():
        rc = lib.bg_rk_solve_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(d_full),
                                    C.byref(launch), sh)
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

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        run_once()
    torch.cuda.synchronize()

    log(f"[bench] timing {args.steps} steps")
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        run_once()
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in ev]
    ms_per_step = wall * 1000.0 / args.steps
    log(f"[bench] ms/step {ms_per_step:.3f} (events {np.mean(kernel_ms):.3f})")

    st = status.cpu().numpy()
    cnt = counts.cpu().numpy().astype(np.int64)
    n_acc = int(cnt[:, 0].sum())
    n_rej = int(cnt[:, 1].sum())
    attempts = n_acc + n_rej
    n_fail = int(np.count_nonzero(st))
    per_inst = cnt[:, 0] + cnt[:, 1]
    solves_per_s = n / (ms_per_step / 1000.0)
    avg_kernel_s = float(np.mean(kernel_ms)) / 1000.0
    flops = attempts * FLOPS_PER_ATTEMPT
    achieved_tflops = flops / avg_kernel_s / 1e12

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
            "parallelism": "one lane per instance, persistent grid, wave-ballot refill",
        },
        "accepted_steps_per_s": n_acc / (ms_per_step / 1000.0),
        "attempts_per_launch": attempts,
        "attempts_per_instance": {"mean": float(per_inst.mean()),
                                  "p99": float(np.percentile(per_inst, 99)),
                                  "max": int(per_inst.max())},
        "failed_instances": n_fail,
        "roofline": {
            "bound": "mfma",
            "compute": "fp64 (MI355X fp64 vector peak = fp64 matrix peak = 78.6 TFLOP/s; "
                       "this kernel issues FP64 VALU fma)",
            "achieved": achieved_tflops,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP64_TFLOPS,
            "traffic": MEASURED_TRAFFIC_BYTES,
            "kernel": "rk_ensemble_kernel<RhsLorenz, Tableau<Fehlberg45,false>>",
            "flops_per_attempt": FLOPS_PER_ATTEMPT,
            "avg_kernel_ms": float(np.mean(kernel_ms)),
            "algorithmic_hbm_bytes": n * BYTES_PER_INSTANCE,
        },
    }

    # plain per-instance path on a subset of the same workload
    if not args.no_plain:
        m = 64
        sub = W.lorenz_batched(m)
        ys, ps = T(sub["y0"]), T(sub["params"])
        te, tl = T(sub["t_end"]), T(sub["tol"])
        to = torch.zeros(m, dtype=torch.float64, device=dev)
        yo = torch.zeros((3, m), dtype=torch.float64, device=dev)
        so = torch.zeros(m, dtype=torch.int32, device=dev)
        co = torch.zeros((m, 2), dtype=torch.int32, device=dev)
        ds = L.EnsembleDesc(n=m, dim=3, n_params=3, y0=ys.data_ptr(), params=ps.data_ptr(),
                            t_end=te.data_ptr(), tol=tl.data_ptr(), dt_min_s=w["dt_min"],
                            dt_max_s=w["dt_max"], t_out=to.data_ptr(), y_out=yo.data_ptr(),
                            status=so.data_ptr(), counts=co.data_ptr())
        log("[bench] plain per-instance path, 64 instances")
        torch.cuda.synchronize()
        tp = time.perf_counter()
        rc = lib.bg_rk_solve_plain_device(L.RK45, L.RHS["lorenz"], L.INTENDED, C.byref(ds), sh)
        torch.cuda.synchronize()
        tp = time.perf_counter() - tp
        if rc != 0:
            raise RuntimeError(f"plain path returned {rc}")
        plain_sps = m / tp
        # same instances in the batched result must be identical
        same = bool(np.array_equal(yo.cpu().numpy(), y_out[:, :m].cpu().numpy()))
        result["plain_gpu"] = {"solves_per_s": plain_sps, "sample": f"first {m} instances",
                               "identical_to_batched": same}
        result["speedup_vs_plain"] = solves_per_s / plain_sps

    if not args.no_cpu_baseline:
        import oracle as O
        threads = os.cpu_count() or 1
        threads = min(threads, 16)
        log(f"[bench] CPU oracle baseline ({threads} threads, ~{args.cpu_seconds}s)")
        done = 0
        chunk = 512
        tc = time.perf_counter()
        while time.perf_counter() - tc < args.cpu_seconds and done < n:
            s = W.lorenz_batched(chunk, first=done)
            O.solve_ensemble(O.RK45, "lorenz", s["y0"], s["params"], s["t_end"], s["tol"],
                             dt_min=w["dt_min"], dt_max=w["dt_max"], semantics=O.INTENDED,
                             profile=O.PROFILE_BACON, threads=threads)
            done += chunk
        tc = time.perf_counter() - tc
        result["cpu_baseline"] = {
            "value": done / tc, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"first {done} instances of the same workload, C oracle "
                      f"(bacon arithmetic, -O2), OpenMP over {threads} threads, {tc:.1f} s",
        }
        result["speedup_vs_cpu"] = solves_per_s / (done / tc)

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
