"""Learned warm start of the 2D Euler and shallow-water closures on the GPU:
the feature map and the linearised start against their definitions (oracle
DS / DU), the gain of a trained network on held-out cells with uniform orders
(wave kernel, the selection fused into it; workgroup kernel for the
dam-break closure, n = 105), and the same under per-cell orders (nested
total degrees, momentsolver.cpp:88,104-117), where the moments past a cell's
order enter the network as zero and its duals there stay zero."""
import numpy as np
import pytest

from uqcreator_amd import native

pytestmark = pytest.mark.gpu


def _cells(oracle, b, n, seed, nt=None, spread=0.02):
    """Duals near the plain start of random states and their moments u."""
    rng = np.random.default_rng(seed)
    c = oracle.closure(oracle.EULER2D, eps=1e-9, safe=1)
    lam_star = np.zeros((n, b.M, 4))
    for j in range(n):
        rho, vx, vy, p = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2)
        lam_star[j, 0] = oracle.DS(c, np.array([rho, rho * vx, rho * vy, p / 0.4 + 0.5 * rho * (vx * vx + vy * vy)]))
        k = b.M if nt is None else nt[j]
        lam_star[j, 1:k] = rng.normal(0, spread, (k - 1, 4)) * np.abs(lam_star[j, 0])
    uq = oracle.eval_uq(c, b.phi, lam_star, nt_cell=nt)
    u = oracle.reconstruct(4, b.phiw, np.ascontiguousarray(uq), nt_cell=nt)
    lam0 = np.zeros_like(lam_star)
    for j in range(n):
        lam0[j, 0] = oracle.DS(c, u[j, 0])
    return c, u, lam_star, lam0


def _scales(u):
    rho, mx, my, E = u[0]
    T = 0.4 * (E - 0.5 * (mx * mx + my * my) / rho) / rho
    return np.array([rho, rho * np.sqrt(T), rho * np.sqrt(T), rho * T])


def _linear_start(oracle, c, u):
    """lambda_0 = DS(u_0), lambda_i = DU(DS(u_0))^{-1} u_i: one full Newton
    step from the plain start, in closed form (orthonormal basis)."""
    ds = oracle.DS(c, u[0])
    J = oracle.DU(c, ds)
    out = np.zeros_like(u)
    out[0] = ds
    for i in range(1, u.shape[0]):
        out[i] = np.linalg.solve(J, u[i])
    return out


def test_euler2d_features_and_linear_start(oracle):
    b = native.Basis(2, [3], 5)  # M = 10, Q = 25: the 2D Euler loop's closure
    c, u, lam_star, lam0 = _cells(oracle, b, 256, seed=3)
    cl = native.GpuClosure(native.EULER2D, b, epsilon=1e-9, safe_line_search=1)
    net = native.WarmStartNet(cl, hidden=64, seed=2)
    assert net.dims[0] == 64 and net.dims[-1] == 64
    du, dl = native.DeviceArray.from_host(u), native.DeviceArray.from_host(lam_star)
    x, t = net.features(du, dl, standardise=False)
    x, t = x.download(), t.download()
    print("case: Euler2D features 256 cells", flush=True)
    for j in (0, 101, 255):
        D = _scales(u[j])
        np.testing.assert_allclose(x[j, :40], (u[j] / D).ravel(), rtol=2e-7, atol=1e-7)
        assert not x[j, 40:].any()
        lin = _linear_start(oracle, c, u[j])
        np.testing.assert_allclose(t[j, :40], ((lam_star[j] - lin) * D / u[j, 0, 0]).ravel(), rtol=2e-5, atol=2e-6)
    # a zero network predicts the linearised start itself
    p, nrm = net.get_params()
    net.set_params(np.zeros_like(p), nrm)
    lp = net.predict(du).download()
    for j in (0, 101, 255):
        np.testing.assert_allclose(lp[j], _linear_start(oracle, c, u[j]), rtol=1e-11, atol=1e-13)


def _train(net, u, lam):
    losses = [net.train(u, lam, batch=1024, steps=400, lr=lr, seed=i + 1) for i, lr in enumerate((1e-3, 3e-4, 1e-4))]
    print("mean losses", losses, flush=True)
    return losses


@pytest.mark.parametrize("precision", [native.FP64, native.MIXED])
def test_euler2d_trained_warm_start_gain(oracle, precision):
    """Uniform orders (wave kernel, selection fused): the warm start keeps the
    fixed point and saves Newton steps on cells it was not trained on."""
    b = native.Basis(2, [3], 5)
    _, u, lam_star, lam0 = _cells(oracle, b, 1 << 14, seed=11)
    cl = native.GpuClosure(native.EULER2D, b, epsilon=1e-9, safe_line_search=1, precision=precision)
    du = native.DeviceArray.from_host(u)
    dl = native.DeviceArray.from_host(lam0)
    print("case: Euler2D warm start train 16384 pairs, precision", precision, flush=True)
    assert cl.solve(du, dl) == 0
    net = native.WarmStartNet(cl, hidden=128, seed=4)
    losses = _train(net, du, dl)
    assert losses[-1] < losses[0]
    _, ut, ls_t, l0_t = _cells(oracle, b, 4096, seed=12)
    dut = native.DeviceArray.from_host(ut)
    ref = native.DeviceArray.from_host(l0_t)
    st = native.DeviceArray(4096, np.int32)
    it = native.DeviceArray(4096, np.int32)
    assert cl.solve(dut, ref, st, it) == 0
    assert cl.last_kernel() == native.KERNEL_WAVE
    plain_steps, _ = cl.last_closure_stats()
    net.attach()
    warm = native.DeviceArray.from_host(l0_t)
    assert cl.solve(dut, warm, st, it) == 0
    warm_steps, _ = cl.last_closure_stats()
    net.detach()
    assert (st.download() == 0).all()
    scale = np.abs(ls_t).max(axis=1, keepdims=True)
    err_plain = (np.abs(ref.download() - ls_t) / scale).max()
    err_warm = (np.abs(warm.download() - ls_t) / scale).max()
    print("Newton steps plain", plain_steps, "warm", warm_steps, "lambda err", err_plain, err_warm, flush=True)
    assert err_plain < 1e-7 and err_warm < 1e-7
    assert warm_steps < plain_steps


def test_warm_start_under_per_cell_orders(oracle):
    """Per-cell orders (workgroup kernel, selection pass with each cell's
    order): same statuses and duals as the oracle's Newton from the plain
    start to the Newton tolerance, duals past a cell's order stay zero, and
    the learned start saves Newton steps."""
    b = native.Basis(2, [1, 2, 3, 4, 5, 6], 7)  # M <= 28, Q = 49
    ntl = b.ntotal_per_level
    rng = np.random.default_rng(21)
    n_train, n_test = 1 << 13, 1024
    nt_train = ntl[rng.integers(0, len(ntl), n_train)].astype(np.int32)
    nt_test = ntl[rng.integers(0, len(ntl), n_test)].astype(np.int32)
    c, u, lam_star, lam0 = _cells(oracle, b, n_train, seed=22, nt=nt_train, spread=0.01)
    cl = native.GpuClosure(native.EULER2D, b, epsilon=1e-9, safe_line_search=1)
    du = native.DeviceArray.from_host(u)
    dl = native.DeviceArray.from_host(lam0)
    dnt = native.DeviceArray.from_host(nt_train)
    print("case: per-cell orders warm start train", n_train, "pairs", flush=True)
    assert cl.solve(du, dl, nt_cell=dnt) == 0
    lam_tr = dl.download()
    for j in range(0, n_train, 97):  # zero-padded pairs, as the solver leaves them
        assert not lam_tr[j, nt_train[j]:].any() and not u[j, nt_train[j]:].any()
    net = native.WarmStartNet(cl, hidden=128, seed=6)
    _train(net, du, dl)
    _, ut, ls_t, l0_t = _cells(oracle, b, n_test, seed=23, nt=nt_test, spread=0.01)
    plain = l0_t.copy()
    st_p, it_p = cl.solve_host(ut, plain, nt_cell=nt_test)
    assert cl.last_kernel() == native.KERNEL_BLOCK
    plain_steps = int(it_p.sum())
    net.attach()
    warm = l0_t.copy()
    st_w, it_w = cl.solve_host(ut, warm, nt_cell=nt_test)
    net.detach()
    warm_steps = int(it_w.sum())
    lam_o = l0_t.copy()
    st_o, it_o = oracle.solve_closure(c, b.phi, b.phiw, ut, lam_o, nt_cell=nt_test)
    assert (st_w == 0).all() and (st_p == 0).all() and (st_o == 0).all()
    scale = np.abs(ls_t).max(axis=1, keepdims=True)
    err = (np.abs(warm - lam_o) / scale).max()
    print("Newton steps plain", plain_steps, "warm", warm_steps, "oracle", int(it_o.sum()), "max rel err", err,
          flush=True)
    # the learned start changes the Newton path: the result is the oracle's
    # fixed point to the Newton tolerance (||grad L|| < 1e-9, checked by the
    # oracle's gradient), not iterate for iterate
    assert err < 1e-7
    for j in range(n_test):
        assert not warm[j, nt_test[j]:].any()
        k = int(nt_test[j])
        gn = np.linalg.norm(oracle.gradient(c, b.phi, b.phiw, np.ascontiguousarray(warm[j]), ut[j], nt=k))
        assert gn < 2e-9, (j, gn)
    assert warm_steps < plain_steps


def _sw_cells(oracle, b, n, seed, barrier, spread=0.02):
    rng = np.random.default_rng(seed)
    okind = oracle.SW2D_BARRIER if barrier else oracle.SW2D
    c = oracle.closure(okind, eps=1e-8, safe=1, barrier=barrier)
    lam_star = np.zeros((n, b.M, 3))
    for j in range(n):
        h, vx, vy = rng.uniform(1, 3), rng.uniform(-1, 1), rng.uniform(-1, 1)
        lam_star[j, 0] = oracle.DS(c, np.array([h, h * vx, h * vy]))
        lam_star[j, 1:] = rng.normal(0, spread, (b.M - 1, 3)) * np.abs(lam_star[j, 0])
    uq = oracle.eval_uq(c, b.phi, lam_star)
    u = oracle.reconstruct(3, b.phiw, np.ascontiguousarray(uq))
    lam0 = np.zeros_like(lam_star)
    for j in range(n):
        lam0[j, 0] = oracle.DS(c, u[j, 0])
    return c, u, lam_star, lam0


SW_CASES = [(native.SHALLOW_WATER2D, 0.0, (2, [3], 5)), (native.SHALLOW_WATER2D_BARRIER, 0.01, (3, [4], 5))]


@pytest.mark.parametrize("kind,barrier,basis", SW_CASES)
def test_shallow_water_features_and_linear_start(oracle, kind, barrier, basis):
    """Scales (h0, h0 a0, h0 a0), a0 = sqrt(g h0); targets times (1/a0^2,
    1/a0, 1/a0); a zero network predicts the linearised start DS(u_0),
    DU(DS(u_0))^{-1} u_i from the oracle's DS and DU."""
    b = native.Basis(*basis)
    c, u, lam_star, lam0 = _sw_cells(oracle, b, 256, seed=41, barrier=barrier)
    cl = native.GpuClosure(kind, b, barrier=barrier, epsilon=1e-8, safe_line_search=1)
    net = native.WarmStartNet(cl, hidden=64, seed=3)
    nf = 3 * b.M
    du, dl = native.DeviceArray.from_host(u), native.DeviceArray.from_host(lam_star)
    x, t = net.features(du, dl, standardise=False)
    x, t = x.download(), t.download()
    print("case: shallow-water features kind", kind, "M", b.M, flush=True)
    for j in (0, 77, 255):
        h = u[j, 0, 0]
        a = np.sqrt(9.81 * h)
        D = np.array([h, h * a, h * a])
        ts = np.array([1 / a ** 2, 1 / a, 1 / a])
        np.testing.assert_allclose(x[j, :nf], (u[j] / D).ravel(), rtol=2e-7, atol=1e-7)
        assert not x[j, nf:].any()
        lin = _linear_start(oracle, c, u[j])
        np.testing.assert_allclose(t[j, :nf], ((lam_star[j] - lin) * ts).ravel(), rtol=2e-5, atol=2e-6)
    p, nrm = net.get_params()
    net.set_params(np.zeros_like(p), nrm)
    lp = net.predict(du).download()
    for j in (0, 77, 255):
        np.testing.assert_allclose(lp[j], _linear_start(oracle, c, u[j]), rtol=1e-11, atol=1e-13)


def test_shallow_water_barrier_warm_start_gain(oracle):
    """The dam-break closure (barrier 0.01, n = 105, 5^3 points, workgroup
    kernel): a trained network keeps the fixed point and saves Newton steps
    on held-out cells."""
    b = native.Basis(3, [4], 5)
    _, u, lam_star, lam0 = _sw_cells(oracle, b, 1 << 13, seed=51, barrier=0.01)
    cl = native.GpuClosure(native.SHALLOW_WATER2D_BARRIER, b, barrier=0.01, epsilon=1e-8, safe_line_search=1)
    du = native.DeviceArray.from_host(u)
    dl = native.DeviceArray.from_host(lam0)
    print("case: shallow-water barrier warm start train 8192 pairs", flush=True)
    assert cl.solve(du, dl) == 0
    net = native.WarmStartNet(cl, hidden=128, seed=8)
    _train(net, du, dl)
    _, ut, ls_t, l0_t = _sw_cells(oracle, b, 1024, seed=52, barrier=0.01)
    plain = l0_t.copy()
    st_p, it_p = cl.solve_host(ut, plain)
    assert cl.last_kernel() == native.KERNEL_BLOCK
    net.attach()
    warm = l0_t.copy()
    st_w, it_w = cl.solve_host(ut, warm)
    net.detach()
    assert (st_p == 0).all() and (st_w == 0).all()
    scale = np.abs(ls_t).max(axis=1, keepdims=True)
    err_plain = (np.abs(plain - ls_t) / scale).max()
    err_warm = (np.abs(warm - ls_t) / scale).max()
    print("Newton steps plain", int(it_p.sum()), "warm", int(it_w.sum()), "lambda err", err_plain, err_warm, flush=True)
    assert err_plain < 1e-7 and err_warm < 1e-7
    assert it_w.sum() < it_p.sum()
