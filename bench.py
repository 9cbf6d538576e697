"""Benchmark of the IPM moment-solver hot path on one MI355X.

Headline workload: the 2D Euler time loop -- transonic flow past an analytic
NACA0012 on a 512 x 128 O-grid (65,536 cells), angle of attack 1.25 + 0.5 xi1
degrees, Mach 0.8 + 0.02 xi2, total degree 3 (M = 10 moments x 4 states = 40
duals per cell), 5x5 Gauss points (Q = 25), Euler entropy closure (Newton to
||grad L||_2 <= 1e-6 on the SI-unit states), Rusanov kinetic flux, CFL 0.4.
One step = one time step of MomentSolver::Solve over every cell: the Newton
closure from the previous step's duals, uQ, the reconstruction, dt, the flux
and the update, the residual (momentsolver.cpp:79-152).  K steps are timed
from the initial state (K = 2000, the configuration's run), replayed from the
captured step graph, with u and lambda resident in HBM.

    python bench.py [--steps K] [--warmup W] [--legs fp64,mixed,mixed_warm_start]
                    [--no-cpu] [--no-c2]

Legs: plain FP64 closure; mixed closure (Hessian on FP32 MFMA, FP64
Cholesky, gradient and accept test); mixed closure with the learned warm start
(a network trained here, outside the timed regions, on (u, lambda) pairs the
FP64 leg's Newton solver produced; the prediction and the per-cell selection
run inside every step).  `value` is the mixed leg.  Each leg reports the
rel. L2 of its mean/variance fields against the FP64 leg after the K steps,
and against the oracle after its first steps.  The closure and advance kernels
are timed with HIP events on the launch stream over the same K steps run again
without the graph (same trajectory), which gives each kernel's roofline.

The same legs also close every cell once from the plain start (the run's
first step, lambda_0 = (DS(u_0), 0, ...)), where the learned warm start
replaces most Newton steps (closure_from_lambda0).

Secondary: the batched-closure benchmark (1D Euler, M = 15, Q = 49, 2^20
cells, one Newton closure of every cell from lambda_0 = (DS(u_0), 0, ...) to
||grad L||_2 <= 1e-10) with the same three legs.

The CPU baselines are the FP64 oracle (OpenMP over cells, dynamic(10)) on
bounded samples of the same workloads, timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from uqcreator_amd.roofline import advance_roofline, closure_roofline, flop_model, rel_l2  # noqa: E402

C3_NI, C3_NJ = 512, 128
C3_EPS = 1e-6
C3_CFL = 0.4


def cpu_threads():
    """The host cores this process may use (the box allots 16 CPUs per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


# --------------------------------------------------------------------------
# C3: the 2D Euler time loop
# --------------------------------------------------------------------------
class C3:
    def __init__(self, native, np):
        from uqcreator_amd import workloads

        self.native, self.np = native, np
        self.b, self.m, self.u0 = workloads.euler2d_c3(C3_NI, C3_NJ)
        self.lam0 = np.zeros_like(self.u0)
        self.lam0[:, 0] = workloads.euler2d_ds(self.u0[:, 0])
        self.prob = native.problem(native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=C3_CFL)
        self.n = self.m.n_cells

    def solver(self, precision, net_params=None):
        from uqcreator_amd import solver

        native = self.native
        cl = native.GpuClosure(native.EULER2D, self.b, epsilon=C3_EPS, precision=precision)
        net = None
        if net_params is not None:
            net = native.WarmStartNet(cl, hidden=256, seed=5)
            net.set_params(*net_params)
            net.attach()
        s = solver.GpuMomentSolver(cl, self.prob, self.m)
        return cl, s, net

    def reset(self, s):
        s.set_state(self.u0, self.lam0)
        s.t, s.steps, s.newton_steps = 0.0, 0, 0

    def leg(self, name, precision, args, net_params=None):
        native, np = self.native, self.np
        cl, s, net = self.solver(precision, net_params)
        # warm-up: W steps from the initial state (captures the step graph)
        self.reset(s)
        s.run_device(1e3, max_steps=args.warmup, use_graph=not args.no_graph)
        self.reset(s)
        native.synchronize()
        t0 = time.perf_counter()
        s.run_device(1e3, max_steps=args.steps, use_graph=not args.no_graph)
        native.synchronize()
        elapsed = time.perf_counter() - t0
        assert s.steps == args.steps, (s.steps, args.steps)
        newton, grads = cl.last_closure_stats()
        factors, lagged = cl.last_closure_factorizations()
        mv = s.mean_var()
        t_end = s.t
        # the same K steps again without the graph: HIP events around every
        # step's closure and advance kernels; Newton steps per cell sampled
        self.reset(s)
        cl.set_kernel_timing(True)
        hist = {}
        chunk = max(1, args.steps // 10)
        done = 0
        while done < args.steps:
            k = min(chunk, args.steps - done)
            s.run_device(1e3, max_steps=k, use_graph=False)
            done += k
            it = s.iters.download()
            hist[done] = np.bincount(np.minimum(it, 5), minlength=6).tolist()  # 0..4 and >= 5
        closure_ms, advance_ms, timed_steps = cl.kernel_times()
        cl.set_kernel_timing(False)
        rerun_same = bool(np.array_equal(s.mean_var(), mv))
        st = s.status.download()
        fm = flop_model(4, self.b.M, self.b.Q)
        S, M, Q = 4, self.b.M, self.b.Q
        # the closure's timing window: the residual check (from the last
        # step's moments after the run's first step), the Newton kernel and,
        # in the mixed time loop with kept inverses, the inverse kernel
        kname = ("closure_recheck_kernel<2,10,25> (closure_check_kernel<2,10,25> in the first step) + " +
                 ("newton_wave_kernel<2,10,25,2> + factor_inverse_kernel<40>" if lagged > 0 else
                  "newton_wave_kernel<2,10,25,%d>" % precision))
        roof_c = closure_roofline(kname, closure_ms / timed_steps, newton / args.steps, grads / args.steps, fm,
                                  precision == native.MIXED, factors / args.steps, lagged / args.steps)
        adv_ms = advance_ms / timed_steps
        roof_a = advance_roofline("advance_kernel<2>", adv_ms, S, M, Q, self.m.nbr.shape[1], 2, self.n)
        out = {
            "value": self.n * args.steps / elapsed,
            "ms_per_step": 1e3 * elapsed / args.steps,
            "t_end": t_end,
            "newton_steps_per_cell_step": newton / (self.n * args.steps),
            "factorizations_per_cell_step": factors / (self.n * args.steps),
            "lagged_steps_per_cell_step": lagged / (self.n * args.steps),
            "gradient_evals_per_cell_step": grads / (self.n * args.steps),
            "newton_histogram_at_step": hist,
            "status_counts": np.bincount(st, minlength=5).tolist(),
            "kernel_ms_per_step": {"closure": closure_ms / timed_steps, "advance": adv_ms,
                                   "steps_timed": timed_steps,
                                   "rest_of_step_in_timed_run": 1e3 * elapsed / args.steps
                                   - closure_ms / timed_steps - adv_ms},
            "rerun_without_graph_identical": rerun_same,
            "roofline": {"closure": roof_c, "advance": roof_a},
        }
        if net is not None:
            net.detach()
        return out, mv

    def cold_closure(self, precision, reps, net_params=None):
        """One closure of every cell from the plain start, the run's first
        step (lambda_0 = (DS(u_0), 0, ...), momentsolver.cpp:463-523 then
        closure.cpp:133-193), the warm start's prediction and selection
        included: wall time around the call, best of `reps`."""
        native, np = self.native, self.np
        cl, _, net = self.solver(precision, net_params)
        n = self.n
        du, dl = native.DeviceArray.from_host(self.u0), native.DeviceArray.from_host(self.lam0)
        st, it = native.DeviceArray(n, np.int32), native.DeviceArray(n, np.int32)
        cl.solve(du, dl, st, it)  # warm-up
        best, nfail = None, 0
        for _ in range(reps):
            dl.upload(self.lam0)
            native.synchronize()
            t0 = time.perf_counter()
            nfail = cl.solve(du, dl, st, it)
            native.synchronize()
            wall = time.perf_counter() - t0
            best = wall if best is None else min(best, wall)
        out = {"cells_per_s": n / best, "wall_ms": 1e3 * best, "failed": int(nfail),
               "newton_steps_per_cell": cl.last_closure_stats()[0] / n}
        lam = dl.download()
        if net is not None:
            net.detach()
        return out, lam

    def check_vs_oracle(self, oracle_py, precision, steps, lam_o_mv):
        """E/Var rel L2 against the oracle after the first `steps` steps."""
        native = self.native
        cl, s, _ = self.solver(precision)
        self.reset(s)
        s.run_device(1e3, max_steps=steps)
        return rel_l2(s.mean_var(), lam_o_mv, self.m.area)

    def oracle_run(self, oracle_py, steps, nthreads):
        np = self.np
        c = oracle_py.closure(oracle_py.EULER2D, eps=C3_EPS)
        p = oracle_py.problem(oracle_py.P_EULER2D, flux=oracle_py.FLUX_LF, cfl=C3_CFL)
        m = self.m
        mesh = oracle_py.MeshArrays(m.nbr, m.unit_normal, m.normal, m.area, m.dx_param, m.bc, m.bnormal)
        u, lam = self.u0.copy(), self.lam0.copy()
        t0 = time.perf_counter()
        res = oracle_py.run(c, p, mesh, self.b.phi, self.b.phiw, self.b.xi, u, lam, 1e3, steps, nthreads=nthreads)
        secs = time.perf_counter() - t0
        if res["rc"] != 0 or res["steps"] != steps:
            raise RuntimeError("oracle C3 run failed: %s" % res)
        mv = oracle_py.mean_var(c, self.b.phi, self.b.phiw, lam, nthreads=nthreads)
        return secs, mv, res


def train_c3_warm_start(native, np, c3, args):
    """Pairs (u, lambda) from the plain FP64 Newton solver along the FP64 run:
    snapshots of every cell at several steps."""
    t0 = time.perf_counter()
    cl, s, _ = c3.solver(native.FP64)
    c3.reset(s)
    n = c3.n
    snaps = max(1, args.c3_train_snapshots)
    every = max(1, args.steps // snaps)
    u = native.DeviceArray((n * snaps, c3.b.M, 4))
    lam = native.DeviceArray((n * snaps, c3.b.M, 4))
    for k in range(snaps):
        s.run_device(1e3, max_steps=every)
        # lambda solves the closure for the moments u_scratch reconstructed from it
        u.copy_from(s.u_scratch, dst_offset=k * n)
        lam.copy_from(s.lam, dst_offset=k * n)
    net = native.WarmStartNet(cl, hidden=256, seed=5)
    steps = args.c3_train_steps // 4
    losses = [net.train(u, lam, batch=1024, steps=steps, lr=lr, seed=i + 1)
              for i, lr in enumerate((1e-3, 5e-4, 2.5e-4, 1e-4))]
    native.synchronize()
    info = {"pairs": n * snaps, "snapshots_every_steps": every, "adam_steps": 4 * steps, "batch": 1024,
            "mean_loss_per_phase": losses, "train_s": time.perf_counter() - t0,
            "network": "features 64 -> 4 x 256 GELU -> 64, BF16 weights, FP32 accumulation"}
    return net.get_params(), info


# --------------------------------------------------------------------------
# C2: the batched closure
# --------------------------------------------------------------------------
def c2_leg(native, np, b, u, lam0, lam_star, args, precision, warm=None):
    cl = native.GpuClosure(native.EULER1D, b, epsilon=1e-10, du_exact=1, safe_line_search=1, precision=precision)
    net = None
    if warm is not None:
        net = native.WarmStartNet(cl, hidden=256, seed=5)
        net.set_params(*warm)
        net.attach()
    lam = native.DeviceArray(lam0.shape)
    st = native.DeviceArray(args.c2_cells, np.int32)
    it = native.DeviceArray(args.c2_cells, np.int32)
    for _ in range(args.c2_warmup):
        lam.copy_from(lam0)
        cl.solve(u, lam, st, it)
    native.synchronize()
    kernel_ms, newton, grads, factors, lagged = [], 0, 0, 0, 0
    t0 = time.perf_counter()
    for _ in range(args.c2_steps):
        lam.copy_from(lam0)
        native.synchronize()
        cl.solve(u, lam, st, it)
        kernel_ms.append(cl.last_closure_ms())
        a, g = cl.last_closure_stats()
        f, lg = cl.last_closure_factorizations()
        newton += a
        grads += g
        factors += f
        lagged += lg
    native.synchronize()
    elapsed = time.perf_counter() - t0
    status = st.download()
    lam_h = lam.download()
    scale = np.abs(lam_star).max(axis=1, keepdims=True)
    dmv = native.DeviceArray((args.c2_cells, 6))
    cl.mean_var(lam, dmv)
    fm = flop_model(3, b.M, b.Q)
    kname = "newton_wave_kernel<1,15,49,%d>" % (4 if lagged > 0 else precision)
    leg = {
        "value": args.c2_cells / (elapsed / args.c2_steps),
        "ms_per_step": 1e3 * elapsed / args.c2_steps,
        "newton_steps_per_cell": newton / args.c2_steps / args.c2_cells,
        "factorizations_per_cell": factors / args.c2_steps / args.c2_cells,
        "lagged_steps_per_cell": lagged / args.c2_steps / args.c2_cells,
        "gradient_evals_per_cell": grads / args.c2_steps / args.c2_cells,
        "status_counts": np.bincount(status, minlength=5).tolist(),
        "lambda_rel_err_vs_known": float((np.abs(lam_h - lam_star) / scale).max()),
        "roofline": closure_roofline(kname, float(np.mean(kernel_ms)), newton / args.c2_steps,
                                     grads / args.c2_steps, fm, precision == native.MIXED, factors / args.c2_steps,
                                     lagged / args.c2_steps, reuse="factor"),
    }
    if net is not None:
        net.detach()
    return leg, lam_h, dmv.download()


def train_c2_warm_start(native, np, b, args):
    """Pairs from the plain FP64 Newton solver on cells of another seed."""
    from uqcreator_amd import workloads

    t0 = time.perf_counter()
    u, _, lam = workloads.closure_euler1d(args.c2_train_cells, basis=b, seed=workloads.SEED + 1, device=True)
    cl = native.GpuClosure(native.EULER1D, b, epsilon=1e-10, du_exact=1, safe_line_search=1)
    if cl.solve(u, lam) != 0:
        raise RuntimeError("training pairs: the plain closure failed")
    net = native.WarmStartNet(cl, hidden=256, seed=5)
    losses = [net.train(u, lam, batch=1024, steps=args.c2_train_steps // 4, lr=lr, seed=k + 1)
              for k, lr in enumerate((1e-3, 5e-4, 2.5e-4, 1e-4))]
    native.synchronize()
    info = {"pairs": args.c2_train_cells, "adam_steps": 4 * (args.c2_train_steps // 4), "batch": 1024,
            "mean_loss_per_phase": losses, "train_s": time.perf_counter() - t0,
            "network": "features 64 -> 4 x 256 GELU -> 64, BF16 weights, FP32 accumulation"}
    return net.get_params(), info


def bench_c2(native, np, args, legs_wanted, oracle_py):
    from uqcreator_amd import workloads

    b = workloads.closure_basis()
    u, lam_star, lam0 = workloads.closure_euler1d(args.c2_cells, basis=b, device=True)
    native.synchronize()
    lam_star_h = lam_star.download()
    legs, mvs, warm, info = {}, {}, None, None
    lam_h = None
    for name in legs_wanted:
        prec = native.FP64 if name == "fp64" else native.MIXED
        if name == "mixed_warm_start" and warm is None:
            warm, info = train_c2_warm_start(native, np, b, args)
        print("c2 leg %s" % name, file=sys.stderr, flush=True)
        legs[name], lam_h, mvs[name] = c2_leg(native, np, b, u, lam0, lam_star_h, args, prec,
                                             warm if name == "mixed_warm_start" else None)
    ref = mvs.get("fp64")
    for name in legs:
        legs[name]["mean_var_rel_l2_vs_fp64"] = rel_l2(mvs[name], ref) if ref is not None else None
    out = {"workload": "batched closure, 1D Euler, 2 uncertain dims, total degree 4 (M=15, 45 duals/cell), "
                       "7x7 Gauss (Q=49), 2^20 cells, Newton from (DS(u0), 0, ...) to ||grad||<=1e-10, "
                       "alpha 0.5, SolveClosureSafe",
           "unit": "cell-steps/s (one closure solve of every cell per step)",
           "steps": args.c2_steps, "warmup": args.c2_warmup, "legs": legs,
           "speedup_vs_fp64": {k: v["value"] / legs["fp64"]["value"] for k, v in legs.items()} if "fp64" in legs
           else None,
           "warm_start_training": info}
    if oracle_py is not None:
        ns = min(args.c2_cpu_sample, args.c2_cells)
        us = np.ascontiguousarray(u.download()[:ns])
        ls = np.ascontiguousarray(lam0.download()[:ns])
        nth = cpu_threads()
        c = oracle_py.closure(oracle_py.EULER1D, eps=1e-10, du_exact=1, safe=1)
        t0 = time.perf_counter()
        st_o, _ = oracle_py.solve_closure(c, b.phi, b.phiw, us, ls, nthreads=nth)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": ns / cpu_s, "unit": "cell-steps/s", "cores": nth, "kind": "port",
                               "sample": "first %d of the %d cells, same closure settings, OpenMP dynamic(10) over "
                                         "cells, %.1f s" % (ns, args.c2_cells, cpu_s)}
        out["gpu_vs_oracle_max_rel_lambda"] = float(
            (np.abs(lam_h[:ns] - ls) / np.abs(ls).max(axis=1, keepdims=True)).max())
        out["oracle_status_counts"] = np.bincount(st_o, minlength=5).tolist()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed time steps of the 2D Euler run")
    ap.add_argument("--warmup", type=int, default=20, help="untimed time steps first (then back to the initial state)")
    ap.add_argument("--legs", default="fp64,mixed,mixed_warm_start")
    ap.add_argument("--oracle-steps", type=int, default=200,
                    help="steps of the CPU oracle run: the CPU baseline and the check of every leg")
    ap.add_argument("--cold-reps", type=int, default=5, help="repeats of the closure from lambda_0 (best taken)")
    ap.add_argument("--c3-train-snapshots", type=int, default=8)
    ap.add_argument("--c3-train-steps", type=int, default=19200)
    ap.add_argument("--no-graph", action="store_true",
                    help="enqueue every step from the host instead of replaying the captured graph (for profilers)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="the closure benchmark only")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="one device copy of this many MiB first (a known-size kernel for HBM counters)")
    ap.add_argument("--c2-cells", type=int, default=1 << 20)
    ap.add_argument("--c2-steps", type=int, default=3)
    ap.add_argument("--c2-warmup", type=int, default=1)
    ap.add_argument("--c2-train-cells", type=int, default=1 << 19)
    ap.add_argument("--c2-train-steps", type=int, default=76800)
    ap.add_argument("--c2-cpu-sample", type=int, default=1 << 16)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: it measures one GPU's share of the mesh (the closure is independent "
              "per cell; the sharded sweep is not built); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from uqcreator_amd import native

    legs_wanted = [x for x in args.legs.split(",") if x]
    for x in legs_wanted:
        if x not in ("fp64", "mixed", "mixed_warm_start"):
            raise SystemExit("unknown leg " + x)
    oracle_py = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py

    if args.calibration_copy_mb > 0:
        n = args.calibration_copy_mb << 17
        a, b = native.DeviceArray(n), native.DeviceArray(n)
        b.copy_from(a)
        native.synchronize()
        print("calibration copy: %d MiB device to device" % args.calibration_copy_mb, file=sys.stderr, flush=True)
        del a, b
    if args.no_c3:
        out = {"closure_benchmark": bench_c2(native, np, args, legs_wanted, oracle_py)}
        print(json.dumps(out), flush=True)
        return
    t0 = time.time()
    c3 = C3(native, np)
    build_s = time.time() - t0
    legs, mvs = {}, {}
    warm, train_info = None, None
    for name in legs_wanted:
        prec = native.FP64 if name == "fp64" else native.MIXED
        if name == "mixed_warm_start" and warm is None:
            print("c3 warm-start training", file=sys.stderr, flush=True)
            warm, train_info = train_c3_warm_start(native, np, c3, args)
        print("c3 leg %s" % name, file=sys.stderr, flush=True)
        legs[name], mvs[name] = c3.leg(name, prec, args, warm if name == "mixed_warm_start" else None)
        print("c3 leg %s: %.4g cell-steps/s, %.3f ms/step, closure %.3f ms, advance %.3f ms"
              % (name, legs[name]["value"], legs[name]["ms_per_step"], legs[name]["kernel_ms_per_step"]["closure"],
                 legs[name]["kernel_ms_per_step"]["advance"]), file=sys.stderr, flush=True)
    # the same legs on one closure of every cell from the plain start (the
    # run's first step), where the warm start replaces most Newton steps
    cold, cold_lam = {}, {}
    for name in legs_wanted:
        prec = native.FP64 if name == "fp64" else native.MIXED
        cold[name], cold_lam[name] = c3.cold_closure(prec, args.cold_reps,
                                                     warm if name == "mixed_warm_start" else None)
        print("c3 closure from lambda_0, %s: %.4g cells/s, %.3f ms, %.2f Newton steps per cell"
              % (name, cold[name]["cells_per_s"], cold[name]["wall_ms"], cold[name]["newton_steps_per_cell"]),
              file=sys.stderr, flush=True)
    if "fp64" in cold_lam:
        lref = cold_lam["fp64"]
        scale = np.maximum(np.abs(lref).max(axis=(1, 2)), 1e-300)
        for name in cold:
            cold[name]["lambda_max_rel_diff_vs_fp64"] = float((np.abs(cold_lam[name] - lref).max(axis=(1, 2))
                                                               / scale).max())
            cold[name]["speedup_vs_fp64"] = cold[name]["cells_per_s"] / cold["fp64"]["cells_per_s"]
    ref = mvs.get("fp64")
    for name in legs:
        legs[name]["mean_var_rel_l2_vs_fp64"] = rel_l2(mvs[name], ref, c3.m.area) if ref is not None else None
        legs[name]["mean_var_max_abs_diff_vs_fp64"] = float(np.abs(mvs[name] - ref).max()) if ref is not None else None
        legs[name]["mean_var_sum"] = float(mvs[name].sum())
    head_name = "mixed" if "mixed" in legs else legs_wanted[-1]
    head = legs[head_name]
    prec_head = native.FP64 if head_name == "fp64" else native.MIXED

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
        "dtype": "f64" if prec_head == native.FP64 else "f64+f32-mfma",
        "data": "synthetic",
        "config": {
            "workload": "2D Euler IPM time loop, NACA0012 O-grid 512x128 (65,536 cells), AoA 1.25+0.5 xi1 deg, "
                        "Mach 0.8+0.02 xi2, total degree 3 (M=10 x 4 states = 40 duals/cell), 5x5 Gauss (Q=25), "
                        "Rusanov flux, CFL 0.4, %d steps from the initial state" % args.steps,
            "n_cells": c3.n, "states": 4, "moments": c3.b.M, "quad_points": c3.b.Q, "leg": head_name,
            "newton_epsilon": C3_EPS, "alpha": 1.0,
            "graph": "none: every step enqueued from the host" if args.no_graph else "step captured once, replayed",
        },
        "roofline": head["roofline"]["closure"],
        "roofline_advance": head["roofline"]["advance"],
        "legs": legs,
        "speedup_vs_fp64": {k: v["value"] / legs["fp64"]["value"] for k, v in legs.items()} if "fp64" in legs
        else None,
        "warm_start_training": train_info,
        "closure_from_lambda0": {"workload": "one closure of the 65,536 cells from lambda_0 = (DS(u_0), 0, ...), "
                                             "the run's first step; wall time around the call, best of %d"
                                             % args.cold_reps, "legs": cold},
        "workload_build_s": build_s,
    }

    if oracle_py is not None:
        nth = cpu_threads()
        ks = args.oracle_steps
        print("c3 oracle %d steps on %d threads" % (ks, nth), file=sys.stderr, flush=True)
        cpu_s, mv_o, res_o = c3.oracle_run(oracle_py, ks, nth)
        out["cpu_baseline"] = {
            "value": c3.n * ks / cpu_s, "unit": "cell-steps/s", "cores": nth, "kind": "port",
            "sample": "the same 65,536-cell run, first %d of its time steps (the first with the Newton from "
                      "(DS(u0), 0, ...)), "
                      "OpenMP dynamic(10) over cells, %.1f s" % (ks, cpu_s)}
        out["oracle_check"] = {"steps": ks, "oracle_newton_steps": res_o["newton_iters"]}
        for name in legs:
            prec = native.FP64 if name == "fp64" else native.MIXED
            if name == "mixed_warm_start":
                continue
            out["oracle_check"]["mean_var_rel_l2_%s" % name] = c3.check_vs_oracle(oracle_py, prec, ks, mv_o)
        out["oracle_check"]["tolerance"] = {"fp64": 1e-8, "mixed": 1e-6}

    if not args.no_c2:
        out["closure_benchmark"] = bench_c2(native, np, args, legs_wanted, oracle_py)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
