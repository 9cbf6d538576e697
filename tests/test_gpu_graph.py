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


@pytest.mark.parametrize("kind", ["burgers", "euler2d", "shallow_water"])
def test_fused_dt_equals_the_dt_sweep(kind):
    """The closure kernels that write uQ also reduce ComputeDt over their cells
    (momentsolver.cpp:127-133); the step's dt must equal, bit for bit, the dt
    sweep (uqk_compute_dt) run afterwards on the uQ the step wrote -- for the
    wavefront kernel (Burgers, Euler 2D) and the workgroup kernel (shallow
    water, n = 105)."""
    from uqcreator_amd import mesh as meshes
    if kind == "burgers":
        b, m, u0, lam0, cl, prob = _burgers()
    elif kind == "euler2d":
        b, m, u0 = workloads.euler2d_c3(64, 16)
        lam0 = np.zeros_like(u0)
        lam0[:, 0] = workloads.euler2d_ds(u0[:, 0])
        cl = native.GpuClosure(native.EULER2D, b, epsilon=1e-6)
        prob = native.problem(native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.4)
    else:
        b, m, u0 = workloads.sw2d_c5(16)
        lam0 = np.zeros_like(u0)
        h = u0[:, 0, 0]
        lam0[:, 0, 0] = 9.81 * h - 0.01 / h  # DS of water at rest with the log barrier
        cl = native.GpuClosure(native.SHALLOW_WATER2D_BARRIER, b, barrier=0.01, epsilon=1e-8, safe_line_search=1)
        prob = native.problem(native.SHALLOW_WATER2D_PROBLEM, cfl=0.45, friction=0.01, friction_dim=2)
    print("case: fused dt", kind, flush=True)
    s = solver.GpuMomentSolver(cl, prob, m)
    s.set_state(u0, lam0)
    for _ in range(3):
        dt, _ = s.step()
        assert cl.last_kernel() in (native.KERNEL_WAVE, native.KERNEL_BLOCK)
        assert dt == cl.compute_dt(prob, s.dmesh, s.uq)


def test_graph_recaptured_when_the_step_changes():
    """A captured step replays only while what it launches is unchanged:
    attaching or detaching the warm-start network, or changing the launch
    shape, must re-capture.  The same solver buffers are reused so the
    argument key alone would match.  The network has zero weights, so it
    predicts the linearised start (one closed-form Newton step from
    (DS(u_0), 0, ...)) and saves Newton steps on the first time step."""
    b, m, u0 = workloads.euler2d_c3(32, 8)
    lam0 = np.zeros_like(u0)
    lam0[:, 0] = workloads.euler2d_ds(u0[:, 0])
    prob = native.problem(native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.4)
    cl = native.GpuClosure(native.EULER2D, b, epsilon=1e-6)
    net = native.WarmStartNet(cl, hidden=64, seed=1)
    net.set_params(np.zeros(net.n_params, np.float32))
    s = solver.GpuMomentSolver(cl, prob, m)

    def run(graph, attached):
        if attached:
            net.attach()
        else:
            net.detach()
        s.set_state(u0, lam0)
        s.t, s.steps, s.newton_steps = 0.0, 0, 0
        s.run_device(1e3, max_steps=2, use_graph=graph)
        return s.lam.download(), s.u.download(), s.newton_steps

    print("case: graph recapture on warm start / shape changes", flush=True)
    a_host = run(False, True)
    a_graph = run(True, True)
    d_graph = run(True, False)
    d_host = run(False, False)
    for x, y in ((a_graph, a_host), (d_graph, d_host)):
        np.testing.assert_array_equal(x[0], y[0])
        np.testing.assert_array_equal(x[1], y[1])
        assert x[2] == y[2]
    print("Newton steps with the warm start", a_host[2], "without", d_host[2], flush=True)
    assert a_host[2] < d_host[2]
    a2 = run(True, True)
    np.testing.assert_array_equal(a2[0], a_host[0])
    cl.set_launch_shape(1, 1)
    d2 = run(True, False)
    np.testing.assert_array_equal(d2[0], d_host[0])
    assert d2[2] == d_host[2]
