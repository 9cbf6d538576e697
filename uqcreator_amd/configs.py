"""Throughput of the configurations besides the bench workload, one JSON line
each:

    python -m uqcreator_amd.configs c3 [--steps N]        2D Euler NACA 512x128, FP64 and mixed
    python -m uqcreator_amd.configs c3-adaptive           the same with refinement levels 1/2/3 (FP64, mixed)
    python -m uqcreator_amd.configs c4 [--steps N]        per-cell orders 1..6: the closure from
                                                          lambda_0, then N steps of the loop
    python -m uqcreator_amd.configs c5 [--steps N]        shallow-water dam break 1024^2

c4 and c5 run three legs each -- plain FP64, mixed, mixed with the learned
warm start -- and train the warm start first on (u, lambda) pairs from the
plain FP64 run (--pairs, 5M by default: snapshots of every cell), outside
the timed regions.  Each leg reports cell-steps/s over the timed steps (the
captured step graph), the closure and advance kernel times of further steps
timed with HIP events, the advance against HBM, Newton steps, and the
mean/variance rel. L2 against the FP64 leg.  Inputs are built in HBM before
the timed regions; c3 runs through uqk_run_config (the input file).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import native, roofline, solver, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
INPUT = os.path.join(HERE, "input")


def _run_input(name, repl, tag):
    src = open(os.path.join(INPUT, name)).read()
    for a, b in repl:
        if a not in src:
            raise RuntimeError("%s: no '%s'" % (name, a))
        src = src.replace(a, b)
    fd, path = tempfile.mkstemp(suffix=".toml", dir=INPUT)
    try:
        with os.fdopen(fd, "w") as f:
            f.write(src)
        print("case:", tag, file=sys.stderr, flush=True)
        rep = native.run_config(path)
    finally:
        os.unlink(path)
    rep["case"] = tag
    print(json.dumps(rep), flush=True)
    return rep


def c3(steps):
    for prec in ("fp64", "mixed"):
        _run_input("euler2d_naca.toml", [("maxSteps = 2000", "maxSteps = %d" % steps),
                                         ('precision = "fp64"', 'precision = "%s"' % prec)], "c3_%s_%d" % (prec, steps))


def c3_adaptive():
    for prec in ("fp64", "mixed"):
        _run_input("euler2d_naca_adaptive.toml", [('precision = "fp64"', 'precision = "%s"' % prec)],
                   "c3_adaptive_%s_2000" % prec)


def _kernel_times(cl, s, steps):
    """Closure and advance kernel ms per step over `steps` further steps
    enqueued without the graph, HIP events around each launch."""
    cl.set_kernel_timing(True)
    s.run_device(1e3, max_steps=steps, use_graph=False)
    closure_ms, advance_ms, n = cl.kernel_times()
    cl.set_kernel_timing(False)
    return closure_ms / max(n, 1), advance_ms / max(n, 1)


def _pairs(cl, prob, m, u, lam0, nt, pairs, every):
    """(u, lambda) pairs of the plain run: every `every` steps a snapshot of
    every cell (lambda solves the closure for the moments u_scratch
    reconstructed from it; moments past a cell's order zeroed)."""
    n = m.n_cells
    S, M = cl.S, cl.M
    snaps = max(1, -(-pairs // n))
    s = solver.GpuMomentSolver(cl, prob, m, nt_cell=nt)
    s.set_state(u, lam0)
    du = native.DeviceArray((n * snaps, M, S))
    dl = native.DeviceArray((n * snaps, M, S))
    mask = None if nt is None else (np.arange(M)[None, :] >= np.asarray(nt)[:, None])
    for k in range(snaps):
        s.run_device(1e3, max_steps=every)
        if mask is None:
            du.copy_from(s.u_scratch, dst_offset=k * n)
        else:
            uh = s.u_scratch.download()
            uh[mask] = 0.0
            du.copy_from(native.DeviceArray.from_host(uh), dst_offset=k * n)
        dl.copy_from(s.lam, dst_offset=k * n)
    return du, dl, n * snaps


def _train(cl, du, dl, adam_steps):
    t0 = time.perf_counter()
    net = native.WarmStartNet(cl, hidden=256, seed=5)
    losses = [net.train(du, dl, batch=1024, steps=adam_steps // 4, lr=lr, seed=i + 1)
              for i, lr in enumerate((1e-3, 5e-4, 2.5e-4, 1e-4))]
    native.synchronize()
    return net.get_params(), {"adam_steps": adam_steps, "batch": 1024, "mean_loss_per_phase": losses,
                              "train_s": time.perf_counter() - t0,
                              "network": "features %d -> 4 x 256 GELU -> %d, BF16 weights, FP32 accumulation"
                                         % (net.dims[0], net.dims[-1])}


def _legs(make_closure, prob, m, u, lam0, nt, steps, params, closure_from_lam0, kernels):
    """The three legs on one workload; returns {leg: report}."""
    n = m.n_cells
    legs, ref_mv, ref_lam = {}, None, None
    for leg, prec, warm in (("fp64", native.FP64, False), ("mixed", native.MIXED, False),
                            ("mixed_warm_start", native.MIXED, True)):
        cl = make_closure(prec)
        net = None
        if warm:
            net = native.WarmStartNet(cl, hidden=256, seed=5)
            net.set_params(*params)
            net.attach()
        r = {}
        S, M, Q = cl.S, cl.M, cl.Q
        if closure_from_lam0:
            # one closure of every cell from the plain start (warm: prediction
            # and selection included), wall time around the call
            du, dl = native.DeviceArray.from_host(u), native.DeviceArray.from_host(lam0)
            dnt = None if nt is None else native.DeviceArray.from_host(nt)
            st, it = native.DeviceArray(n, np.int32), native.DeviceArray(n, np.int32)
            cl.solve(du, dl, st, it, nt_cell=dnt)
            dl.upload(lam0)
            native.synchronize()
            t0 = time.perf_counter()
            nfail = cl.solve(du, dl, st, it, nt_cell=dnt)
            native.synchronize()
            wall = time.perf_counter() - t0
            lam = dl.download()
            if ref_lam is None:
                ref_lam = lam
            scale = np.maximum(np.abs(ref_lam).max(axis=(1, 2)), 1e-300)
            r["closure_from_lambda0"] = {
                "cells_per_s": n / wall, "wall_ms": 1e3 * wall, "kernel_ms": cl.last_closure_ms(),
                "failed": int(nfail), "newton_steps_per_cell": cl.last_closure_stats()[0] / n,
                "lambda_max_rel_diff_vs_fp64": float((np.abs(lam - ref_lam).max(axis=(1, 2)) / scale).max())}
            print("case:", leg, "closure from lambda_0", r["closure_from_lambda0"], file=sys.stderr, flush=True)
        s = solver.GpuMomentSolver(cl, prob, m, nt_cell=nt)
        s.set_state(u, lam0)
        s.run_device(1e3, max_steps=2)  # capture and warm-up
        native.synchronize()
        t0 = time.perf_counter()
        s.run_device(1e3, max_steps=steps)
        native.synchronize()
        wall = time.perf_counter() - t0
        newton, grads = cl.last_closure_stats()
        factors, lagged = cl.last_closure_factorizations()
        mv = s.mean_var()
        if ref_mv is None:
            ref_mv = mv
        cms, ams = _kernel_times(cl, s, max(1, min(steps, 10)))
        r.update({
            "value": n * steps / wall, "unit": "cell-steps/s", "ms_per_step": 1e3 * wall / steps, "steps": steps,
            "newton_steps_per_cell_step": newton / (n * steps), "factorizations_per_cell_step": factors / (n * steps),
            "lagged_steps_per_cell_step": lagged / (n * steps), "gradient_evals_per_cell_step": grads / (n * steps),
            "mean_var_rel_l2_vs_fp64": roofline.rel_l2(mv, ref_mv, m.area),
            "kernel_ms_per_step": {"closure": cms, "advance": ams, "closure_kernel": cl.last_kernel()},
            "roofline_advance": roofline.advance_roofline(kernels["advance"], ams, S, M, Q, m.nbr.shape[1], 2, n)})
        if nt is None:
            fm = roofline.flop_model(S, M, Q)
            block = cl.last_kernel() == native.KERNEL_BLOCK
            # (the workgroup kernel's mixed time loop is its kept-factor instantiation, <KIND, 2>)
            r["roofline_closure"] = roofline.closure_roofline(
                kernels["closure"] % (2 if block and prec == native.MIXED else prec), cms, newton / steps,
                grads / steps, fm, prec == native.MIXED, factors / steps, lagged / steps)
        print("case:", leg, "loop", {k: r[k] for k in ("value", "ms_per_step", "newton_steps_per_cell_step",
                                                       "mean_var_rel_l2_vs_fp64")}, file=sys.stderr, flush=True)
        if net is not None:
            net.detach()
        legs[leg] = r
    return legs


def c4(steps, pairs=5_000_000, adam_steps=19200):
    """Per-cell orders 1..6 on the 2D Euler NACA setup (7x7 points, n up to
    112: the workgroup kernel)."""
    b, m, u, lam0, nt = workloads.euler2d_c4()
    prob = native.problem(native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.4)

    def make(prec):
        return native.GpuClosure(native.EULER2D, b, epsilon=1e-6, precision=prec)

    print("case: c4 training pairs", pairs, file=sys.stderr, flush=True)
    cl = make(native.FP64)
    du, dl, npairs = _pairs(cl, prob, m, u, lam0, nt, pairs, every=2)
    params, info = _train(cl, du, dl, adam_steps)
    del du, dl
    info["pairs"] = npairs
    legs = _legs(make, prob, m, u, lam0, nt, steps, params, True,
                 {"closure": "newton_block_kernel<2,%d>", "advance": "advance_kernel"})
    out = {"case": "c4", "workload": "2D Euler NACA0012 O-grid 512x128, per-cell orders 1..6 (nTotal 3..28, "
                                     "seeded), 7x7 Gauss, Rusanov flux, CFL 0.4",
           "n_cells": m.n_cells, "warm_start_training": info, "legs": legs,
           "speedup_vs_fp64": {k: v["value"] / legs["fp64"]["value"] for k, v in legs.items()},
           "closure_from_lambda0_speedup_vs_fp64": {
               k: v["closure_from_lambda0"]["cells_per_s"] / legs["fp64"]["closure_from_lambda0"]["cells_per_s"]
               for k, v in legs.items()}}
    print(json.dumps(out), flush=True)
    return out


def c5(steps, pairs=5_000_000, adam_steps=19200):
    """Shallow-water dam break 1024^2 (the input file's IC and mesh)."""
    inf, u, lam0, center = native.config_setup(os.path.join(INPUT, "shallowwater2d_dambreak.toml"))
    from . import mesh as _mesh

    m = _mesh.quad_grid(1024, 1024, -1.0, 1.0, -1.0, 1.0, native.BC_DIRICHLET)
    if not np.allclose(m.center, center, rtol=0, atol=1e-12):
        raise RuntimeError("c5: the mesh differs from the input file's")
    b = native.Basis(3, [4], 5)
    if (inf["M"], inf["Q"]) != (b.M, b.Q):
        raise RuntimeError("c5: the basis differs from the input file's")
    prob = native.problem(native.SHALLOW_WATER2D_PROBLEM, flux=native.FLUX_REFERENCE, cfl=0.45, friction=0.01,
                          friction_dim=2)

    def make(prec):
        return native.GpuClosure(native.SHALLOW_WATER2D_BARRIER, b, barrier=0.01, epsilon=1e-8, precision=prec)

    print("case: c5 training pairs", pairs, file=sys.stderr, flush=True)
    cl = make(native.FP64)
    du, dl, npairs = _pairs(cl, prob, m, u, lam0, None, pairs, every=2)
    params, info = _train(cl, du, dl, adam_steps)
    del du, dl
    info["pairs"] = npairs
    legs = _legs(make, prob, m, u, lam0, None, steps, params, False,
                 {"closure": "newton_block_kernel<4,%d>", "advance": "advance_kernel"})
    out = {"case": "c5", "workload": "shallow-water dam break 1024x1024, barrier 0.01, total degree 4 (35 x 3 = "
                                     "105 duals), 5^3 Gauss, HLL flux, friction 0.01 (1 + xi3), CFL 0.45",
           "n_cells": m.n_cells, "warm_start_training": info, "legs": legs,
           "speedup_vs_fp64": {k: v["value"] / legs["fp64"]["value"] for k, v in legs.items()}}
    print(json.dumps(out), flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=["c3", "c3-adaptive", "c4", "c5"])
    ap.add_argument("--steps", type=int)
    ap.add_argument("--pairs", type=int, default=5_000_000, help="warm-start training pairs (c4, c5)")
    ap.add_argument("--adam-steps", type=int, default=19200)
    a = ap.parse_args(argv)
    if a.config == "c3":
        c3(a.steps or 2000)
    elif a.config == "c3-adaptive":
        c3_adaptive()
    elif a.config == "c4":
        c4(a.steps or 200, a.pairs, a.adam_steps)
    else:
        c5(a.steps or 20, a.pairs, a.adam_steps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
