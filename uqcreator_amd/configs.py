"""Throughput of the configurations besides the bench workload, one JSON line
each:

    python -m uqcreator_amd.configs c3 [--steps N]        2D Euler NACA 512x128, FP64 and mixed
    python -m uqcreator_amd.configs c3-adaptive           the same with refinement levels 1/2/3
    python -m uqcreator_amd.configs c4 [--steps N]        per-cell orders 1..6: closure from
                                                          lambda_0, then N steps of the loop
    python -m uqcreator_amd.configs c5 [--steps N]        shallow-water dam break 1024^2

Inputs are built in HBM before the timed regions; the time loops run through
uqk_run_steps (captured step graph).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import native, solver, workloads

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
    _run_input("euler2d_naca_adaptive.toml", [], "c3_adaptive_2000")


def c5(steps):
    for prec in ("mixed", "fp64"):
        _run_input("shallowwater2d_dambreak.toml", [("maxSteps = 500", "maxSteps = %d" % steps),
                                                    ('precision = "mixed"', 'precision = "%s"' % prec)],
                   "c5_%s_%d" % (prec, steps))


def c4(steps):
    b, m, u, lam0, nt = workloads.euler2d_c4()
    n = m.n_cells
    for prec, tag in ((native.FP64, "fp64"), (native.MIXED, "mixed")):
        cl = native.GpuClosure(native.EULER2D, b, epsilon=1e-6, precision=prec)
        du, dl, dnt = native.DeviceArray.from_host(u), native.DeviceArray.from_host(lam0), native.DeviceArray.from_host(nt)
        st, it = native.DeviceArray(n, np.int32), native.DeviceArray(n, np.int32)
        print("case: c4 closure", tag, file=sys.stderr, flush=True)
        cl.solve(du, dl, st, it, nt_cell=dnt)  # warm-up
        dl.upload(lam0)
        nfail = cl.solve(du, dl, st, it, nt_cell=dnt)
        ms = cl.last_closure_ms()
        print(json.dumps(dict(case="c4_closure_" + tag, n_cells=n, failed=int(nfail), kernel_ms=ms,
                              cells_per_s=n / (ms * 1e-3), kernel=cl.last_kernel(),
                              newton_steps=cl.last_closure_stats()[0])), flush=True)
        print("case: c4 loop", tag, steps, file=sys.stderr, flush=True)
        prob = native.problem(native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.4)
        s = solver.GpuMomentSolver(cl, prob, m, nt_cell=nt)
        s.set_state(u, lam0)
        s.run_device(1e3, 5)  # capture and warm-up
        t0 = time.time()
        s.run_device(1e3, steps)
        t = time.time() - t0
        print(json.dumps(dict(case="c4_loop_%s_%d" % (tag, steps), n_cells=n, steps=steps, wall_s=t,
                              cell_steps_per_s=n * steps / t)), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=["c3", "c3-adaptive", "c4", "c5"])
    ap.add_argument("--steps", type=int)
    a = ap.parse_args(argv)
    if a.config == "c3":
        c3(a.steps or 2000)
    elif a.config == "c3-adaptive":
        c3_adaptive()
    elif a.config == "c4":
        c4(a.steps or 200)
    else:
        c5(a.steps or 20)
    return 0


if __name__ == "__main__":
    sys.exit(main())
