"""Benchmark of the batched IPM closure on one MI355X.

Workload: the batched-closure benchmark -- 1D Euler (gamma 1.4), two uniform
uncertain dimensions, total-degree-4 Legendre basis (M = 15 moments per state,
45 duals per cell), 7x7 Gauss quadrature (Q = 49), 2^20 cells whose moments
come from seeded random duals around a Maxwellian (uqcreator_amd.workloads).
One step = one Newton closure solve of every cell from lambda_0 = (DS(u_0),
0, ...) to ||grad L||_2 <= 1e-10, i.e. one cell-step per cell.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed|both]
                    [--no-warm-start] [--no-cpu]

Legs: plain FP64 closure; mixed closure (FP32 MFMA Hessian, FP64 Cholesky,
gradient and accept test); mixed closure with the learned warm start (a
network trained here, outside the timed region, on Newton-solved pairs of
other cells; the prediction and the per-cell selection run inside the timed
region).  `value` is the last leg run; every leg reports its throughput, its
error against the known duals and the rel. L2 of the mean/variance fields
against the FP64 leg.  Inputs are resident in HBM before the timed region.
The CPU baseline is the FP64 oracle (OpenMP) on a bounded sample of the same
cells, timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# dense peaks of one MI355X (TFLOP/s)
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def flop_model(S, M, Q):
    """Algorithmic flops of one Newton step (symmetric Hessian Q n (n+1),
    Cholesky + solve n^3/3 + 2 n^2) and of one gradient evaluation (4 Q S M)."""
    n = S * M
    return dict(F_H=Q * n * (n + 1), F_C=n ** 3 / 3.0 + 2.0 * n * n, F_g=4.0 * Q * S * M)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def load_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE scaled by a copy kernel of known size)."""
    path = os.path.join(ROOT, "profiles", "closure_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def run_leg(native, np, b, u, lam0, lam_star, args, precision, warm=None):
    cl = native.GpuClosure(native.EULER1D, b, epsilon=1e-10, du_exact=1, safe_line_search=1, precision=precision)
    cl.set_kernel({"auto": native.KERNEL_AUTO, "generic": native.KERNEL_GENERIC, "wave": native.KERNEL_WAVE}[args.kernel])
    net = None
    if warm is not None:
        net = native.WarmStartNet(cl, hidden=256, seed=5)
        net.set_params(*warm)
        net.attach()
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
    status = st.download()
    lam_h = lam.download()
    scale = np.abs(lam_star).max(axis=1, keepdims=True)
    dmv = native.DeviceArray((args.ncell, 6))
    cl.mean_var(lam, dmv)
    mv = dmv.download()
    S, M, Q = 3, b.M, b.Q
    fm = flop_model(S, M, Q)
    newton_per_launch = newton / args.steps
    grads_per_launch = grads / args.steps
    kms = float(np.mean(kernel_ms))
    f_hess = newton_per_launch * fm["F_H"]
    f_fp64 = newton_per_launch * fm["F_C"] + grads_per_launch * fm["F_g"]
    flops = f_hess + f_fp64
    # the time the flops need at the dense peak of the unit each runs on:
    # Hessian on FP64 (plain) or FP32 (mixed) MFMA, the rest FP64
    hess_peak = FP64_PEAK_TFLOPS if precision == native.FP64 else FP32_PEAK_TFLOPS
    t_min = f_hess / (hess_peak * 1e12) + f_fp64 / (FP64_PEAK_TFLOPS * 1e12)
    achieved = flops / (kms * 1e-3) / 1e12
    peak = flops / t_min / 1e12
    kernel = "newton_wave_kernel" if cl.last_kernel() == native.KERNEL_WAVE else "newton_fp64_kernel"
    leg = {
        "value": args.ncell / (elapsed / args.steps),
        "ms_per_step": 1e3 * elapsed / args.steps,
        "closure_kernel_ms": kms,
        "newton_steps_per_cell": newton_per_launch / args.ncell,
        "gradient_evals_per_cell": grads_per_launch / args.ncell,
        "status_counts": np.bincount(status, minlength=5).tolist(),
        "lambda_rel_err_vs_known": float((np.abs(lam_h - lam_star) / scale).max()),
        "roofline": {
            "bound": "mfma",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": load_traffic(kernel),
            "kernel": kernel,
            "kernel_ms": kms,
            "flops_per_launch": flops,
            "hessian_flops_per_launch": f_hess,
            "flop_model": fm,
        },
    }
    if net is not None:
        net.detach()
    return leg, lam_h, mv


def train_warm_start(native, np, b, args):
    """Pairs from the plain FP64 Newton solver on cells of another seed."""
    from uqcreator_amd import workloads

    t0 = time.perf_counter()
    u, _, lam = workloads.closure_euler1d(args.train_cells, basis=b, seed=workloads.SEED + 1, device=True)
    cl = native.GpuClosure(native.EULER1D, b, epsilon=1e-10, du_exact=1, safe_line_search=1)
    if cl.solve(u, lam) != 0:
        raise RuntimeError("training pairs: the plain closure failed")
    net = native.WarmStartNet(cl, hidden=256, seed=5)
    losses = []
    for k, lr in enumerate((1e-3, 5e-4, 2.5e-4, 1e-4)):
        losses.append(net.train(u, lam, batch=1024, steps=args.train_steps // 4, lr=lr, seed=k + 1))
    native.synchronize()
    info = {"pairs": args.train_cells, "adam_steps": 4 * (args.train_steps // 4), "batch": 1024,
            "mean_loss_per_phase": losses, "train_s": time.perf_counter() - t0,
            "network": "features 64 -> 4 x 256 GELU -> 64, BF16 weights, FP32 accumulation"}
    return net.get_params(), info


def rel_l2(a, b):
    import numpy as np

    num = ((a - b) ** 2).sum(0)
    den = (b ** 2).sum(0)
    return float(np.sqrt(num / np.maximum(den, 1e-20 * den.max())).max())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncell", type=int, default=1 << 20)
    ap.add_argument("--precision", choices=["fp64", "mixed", "both"], default="both")
    ap.add_argument("--no-warm-start", action="store_true")
    ap.add_argument("--train-cells", type=int, default=1 << 19)
    ap.add_argument("--train-steps", type=int, default=76800)
    ap.add_argument("--cpu-sample", type=int, default=1 << 17)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel", choices=["auto", "generic", "wave"], default="auto")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the closure is independent per cell and this benchmark measures one "
              "GPU's share; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from uqcreator_amd import native, workloads

    b = workloads.closure_basis()
    t0 = time.time()
    u, lam_star, lam0 = workloads.closure_euler1d(args.ncell, basis=b, device=True)
    native.synchronize()
    gen_s = time.time() - t0
    lam_star_h = lam_star.download()

    legs, mvs = {}, {}
    plan = []
    if args.precision in ("fp64", "both"):
        plan.append(("fp64", native.FP64, False))
    if args.precision in ("mixed", "both"):
        plan.append(("mixed", native.MIXED, False))
        if not args.no_warm_start:
            plan.append(("mixed_warm_start", native.MIXED, True))
    elif args.precision == "fp64" and not args.no_warm_start:
        plan.append(("fp64_warm_start", native.FP64, True))
    warm, train_info = None, None
    lam_h = None
    for name, prec, use_warm in plan:
        if use_warm and warm is None:
            warm, train_info = train_warm_start(native, np, b, args)
        leg, lam_h, mvs[name] = run_leg(native, np, b, u, lam0, lam_star_h, args, prec, warm if use_warm else None)
        legs[name] = leg
        print("leg %s: %.4g cell-steps/s, %.2f Newton steps/cell" % (name, leg["value"], leg["newton_steps_per_cell"]),
              file=sys.stderr, flush=True)
    ref = mvs.get("fp64")
    for name in legs:
        legs[name]["mean_var_rel_l2_vs_fp64"] = rel_l2(mvs[name], ref) if ref is not None else None
    head_name = plan[-1][0]
    head = legs[head_name]

    out = {
        "metric": "IPM cell-steps/sec incl. Newton closure; mixed vs plain FP64 at equal L2 error",
        "value": head["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if plan[-1][1] == native.FP64 else "f64+f32-mfma",
        "data": "synthetic",
        "config": {
            "workload": "batched closure, 1D Euler, 2 uncertain dims, total degree 4 (M=15, 45 duals/cell), "
                        "7x7 Gauss (Q=49), 2^20 cells, Newton to ||grad||<=1e-10",
            "n_cells": args.ncell, "states": 3, "moments": b.M, "quad_points": b.Q,
            "leg": head_name, "alpha": 0.5, "line_search": "SolveClosureSafe",
        },
        "roofline": head["roofline"],
        "legs": legs,
        "speedup_vs_fp64": {k: v["value"] / legs["fp64"]["value"] for k, v in legs.items()} if "fp64" in legs else None,
        "warm_start_training": train_info,
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
        out["cpu_baseline"] = {
            "value": ns / cpu_s,
            "unit": "cell-steps/s",
            "cores": nth,
            "kind": "port",
            "sample": "first %d of the %d cells, same closure settings, OpenMP dynamic(10) over cells, %.1f s"
                      % (ns, args.ncell, cpu_s),
        }
        out["gpu_vs_oracle_max_rel_lambda"] = float(
            (np.abs(lam_h[:ns] - ls) / np.abs(ls).max(axis=1, keepdims=True)).max())
        out["oracle_status_counts"] = np.bincount(st_o, minlength=5).tolist()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
