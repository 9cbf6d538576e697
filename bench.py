"""Benchmark of the batched IPM closure on one MI355X.

Workload: the batched-closure benchmark -- 1D Euler (gamma 1.4), two uniform
uncertain dimensions, total-degree-4 Legendre basis (M = 15 moments per state,
45 duals per cell), 7x7 Gauss quadrature (Q = 49), 2^20 cells whose moments
come from seeded random duals around a Maxwellian (uqcreator_amd.workloads).
One step = one Newton closure solve of every cell from lambda_0 = (DS(u_0),
0, ...) to ||grad L||_2 <= 1e-10, i.e. one cell-step per cell.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed]

Prints one JSON line (see the keys below); inputs are resident in HBM before
the timed region.  The CPU baseline is the FP64 oracle (OpenMP) on a bounded
sample of the same cells, timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic flop model of one Newton step (symmetric Hessian Q n (n+1),
# Cholesky + solve n^3/3 + 2 n^2) and of one gradient evaluation (4 Q S M)
FP64_PEAK_TFLOPS = 78.6


def flop_model(S, M, Q):
    n = S * M
    return dict(F_H=Q * n * (n + 1), F_C=n ** 3 / 3.0 + 2.0 * n * n, F_g=4.0 * Q * S * M)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncell", type=int, default=1 << 20)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-sample", type=int, default=1 << 14)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", choices=["auto", "generic", "wave"], default="auto")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the closure is independent per cell and this build measures one "
              "GPU's share; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from uqcreator_amd import native, workloads

    b = workloads.closure_basis()
    S, M, Q = 3, b.M, b.Q
    t0 = time.time()
    u, lam_star, lam0 = workloads.closure_euler1d(args.ncell, basis=b, device=True)
    native.synchronize()
    gen_s = time.time() - t0
    prec = native.FP64 if args.precision == "fp64" else native.MIXED
    cl = native.GpuClosure(native.EULER1D, b, epsilon=1e-10, du_exact=1, safe_line_search=1, precision=prec)
    cl.set_kernel({"auto": native.KERNEL_AUTO, "generic": native.KERNEL_GENERIC, "wave": native.KERNEL_WAVE}[args.kernel])
    lam = native.DeviceArray(lam0.shape)
    st = native.DeviceArray(args.ncell, np.int32)
    it = native.DeviceArray(args.ncell, np.int32)

    for _ in range(args.warmup):
        lam.copy_from(lam0)
        cl.solve(u, lam, st, it)
    native.synchronize()
    kernel_ms, newton, grads = [], 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lam.copy_from(lam0)
        native.synchronize()
        cl.solve(u, lam, st, it)
        kernel_ms.append(cl.last_closure_ms())
        a, g = cl.last_closure_stats()
        newton += a
        grads += g
    native.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.ncell / (elapsed / args.steps)

    status = st.download()
    iters = it.download()
    lam_h = lam.download()
    lam_star_h = lam_star.download()
    scale = np.abs(lam_star_h).max(axis=1, keepdims=True)
    lam_err = float((np.abs(lam_h - lam_star_h) / scale).max())

    fm = flop_model(S, M, Q)
    newton_per_launch = newton / args.steps
    grads_per_launch = grads / args.steps
    flops = newton_per_launch * (fm["F_H"] + fm["F_C"]) + grads_per_launch * fm["F_g"]
    kms = float(np.mean(kernel_ms))
    achieved = flops / (kms * 1e-3) / 1e12

    out = {
        "metric": "IPM cell-steps/sec incl. Newton closure; mixed vs plain FP64 at equal L2 error",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == native.FP64 else "f64+f32-mfma",
        "data": "synthetic",
        "config": {
            "workload": "batched closure, 1D Euler, 2 uncertain dims, total degree 4 (M=15, 45 duals/cell), "
                        "7x7 Gauss (Q=49), 2^20 cells, Newton to ||grad||<=1e-10",
            "n_cells": args.ncell, "states": S, "moments": M, "quad_points": Q,
            "precision": args.precision, "alpha": 0.5, "line_search": "SolveClosureSafe",
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "kernel": "newton_wave_kernel" if cl.last_kernel() == native.KERNEL_WAVE else "newton_fp64_kernel",
            "kernel_ms": kms,
            "flops_per_launch": flops,
            "flop_model": fm,
        },
        "newton_steps_per_cell": newton_per_launch / args.ncell,
        "gradient_evals_per_cell": grads_per_launch / args.ncell,
        "status_counts": np.bincount(status, minlength=5).tolist(),
        "iterations_p50_max": [float(np.percentile(iters, 50)), int(iters.max())],
        "lambda_rel_err_vs_known": lam_err,
        "workload_build_s": gen_s,
    }

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py

        ns = min(args.cpu_sample, args.ncell)
        us = np.ascontiguousarray(u.download()[:ns])
        ls = np.ascontiguousarray(lam0.download()[:ns])
        nth = cpu_threads()
        c = oracle_py.closure(oracle_py.EULER1D, eps=1e-10, du_exact=1, safe=1)
        t0 = time.perf_counter()
        st_o, it_o = oracle_py.solve_closure(c, b.phi, b.phiw, us, ls, nthreads=nth)
        cpu_s = time.perf_counter() - t0
        lam_gpu = lam_h[:ns]
        out["cpu_baseline"] = {
            "value": ns / cpu_s,
            "unit": "cell-steps/s",
            "cores": nth,
            "kind": "port",
            "sample": "first %d of the %d cells, same closure settings, OpenMP dynamic(10) over cells, %.1f s"
                      % (ns, args.ncell, cpu_s),
        }
        out["gpu_vs_oracle_max_rel_lambda"] = float(
            (np.abs(lam_gpu - ls) / np.abs(ls).max(axis=1, keepdims=True)).max())
        out["oracle_status_counts"] = np.bincount(st_o, minlength=5).tolist()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
