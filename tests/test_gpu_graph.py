"""The device-resident time loop (uqk_run_steps): the step replayed from a
captured hipGraph must reproduce the host-driven loop exactly, replay after
replay, and stop at t_end like momentsolver.cpp:79."""
import numpy as np
import pytest

from uqcreator_amd import native, solver, workloads

pytestmark = pytest.mark.gpu


def _burgers(n=400):
    b, m, u0, lam0 = workloads.burgers_c1(n)
    cl = native.GpuClosure(native.BOUNDED_BARRIER, b, u_minus=2.9, u_plus=12.1, epsilon=1e-10)
    prob = native.problem(native.BURGERS, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.5)
    return b, m, u0, lam0, cl, prob


@pytest.mark.parametrize("use_graph", [True, False])
def test_run_steps_matches_host_loop(use_graph):
    b, m, u0, lam0, cl, prob = _burgers()
    print("case: run_steps graph", use_graph, flush=True)
    ref = solver.GpuMomentSolver(cl, prob, m)
    ref.set_state(u0, lam0)
    ref.run(0.05, 150)
    dev = solver.GpuMomentSolver(cl, prob, m)
    dev.set_state(u0, lam0)
    # three chunks: the graph is captured once and replayed across calls
    for stop in (40, 100, 150):
        dev.run_device(0.05, max_steps=stop - dev.steps, use_graph=use_graph)
    assert dev.steps == ref.steps == 150
    assert dev.t == ref.t
    np.testing.assert_array_equal(dev.u.download(), ref.u.download())
    np.testing.assert_array_equal(dev.lam.download(), ref.lam.download())
    assert dev.newton_steps == ref.newton_steps


def test_run_steps_stops_at_t_end():
    b, m, u0, lam0, cl, prob = _burgers()
    print("case: run_steps t_end", flush=True)
    ref = solver.GpuMomentSolver(cl, prob, m)
    ref.set_state(u0, lam0)
    ref.run(0.01)
    dev = solver.GpuMomentSolver(cl, prob, m)
    dev.set_state(u0, lam0)
    dev.run_device(0.01)
    assert dev.steps == ref.steps and dev.t == ref.t >= 0.01
    np.testing.assert_array_equal(dev.u.download(), ref.u.download())
