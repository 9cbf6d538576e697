"""Learned warm start on the GPU: the BF16 MFMA forward and backward kernels
against the FP64 oracle network (with the device's BF16 rounding points
emulated), the feature/target map against its definition, and what the
trained network gains over the plain Newton start on held-out cells."""
import numpy as np
import pytest

from uqcreator_amd import native, workloads

pytestmark = pytest.mark.gpu


def _closure_c2():
    b = workloads.closure_basis()
    return b, native.GpuClosure(native.EULER1D, b, epsilon=1e-10, safe_line_search=1)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _split(params, dims):
    out, o = [], 0
    for l in range(len(dims) - 1):
        N, K = dims[l + 1], dims[l]
        out.append((params[o:o + N * K].reshape(N, K), params[o + N * K:o + N * K + N]))
        o += N * K + N
    return out


def _random_params(dims, rng, scale=1.0):
    p = []
    nf = 45
    for l in range(len(dims) - 1):
        N, K = dims[l + 1], dims[l]
        W = rng.uniform(-1, 1, (N, K)) * scale * np.sqrt(3.0 / K)
        if l == 0:
            W[:, nf:] = 0.0
        if l == len(dims) - 2:
            W[nf:, :] = 0.0
        p += [W.ravel(), rng.normal(0, 0.1, N) * (np.arange(N) < nf if l == len(dims) - 2 else 1)]
    return np.concatenate(p).astype(np.float32)


def test_forward_and_gradient_match_oracle(oracle):
    b, cl = _closure_c2()
    net = native.WarmStartNet(cl, hidden=256, seed=3)
    dims = net.dims
    assert dims == [64, 256, 256, 256, 256, 64]
    assert net.n_params == oracle.mlp_param_count(dims)
    rng = np.random.default_rng(11)
    params = _random_params(dims, rng)
    net.set_params(params)
    p_dev, _ = net.get_params()
    np.testing.assert_array_equal(p_dev, params)
    rows = 96  # not a multiple of the 64-row tile: exercises the edge guards
    x = np.zeros((rows, dims[0]), np.float32)
    x[:, :45] = rng.normal(0, 1, (rows, 45))
    t = np.zeros((rows, dims[-1]), np.float32)
    t[:, :45] = rng.normal(0, 1, (rows, 45))
    print("case: mlp forward/backward rows", rows, flush=True)
    y = net.forward(native.DeviceArray.from_host(x)).download()
    y_o = oracle.mlp_forward(dims, params.astype(np.float64), x.astype(np.float64), emulate_bf16=True)
    # what remains is FP32 accumulation order, and the BF16 rounding of an
    # activation flipping by one unit where that order moves it across a
    # rounding boundary
    err_emul = _rel(y, y_o)
    # without the BF16 rounding the difference is the BF16 error itself
    err_64 = _rel(y, oracle.mlp_forward(dims, params.astype(np.float64), x.astype(np.float64)))
    print("forward rel err vs emulated BF16 oracle", err_emul, "vs FP64 oracle", err_64, flush=True)
    assert err_emul < 1e-3 and err_emul < 0.2 * err_64 and err_64 < 5e-2
    loss, g = net.loss_grad(native.DeviceArray.from_host(x), native.DeviceArray.from_host(t), 45)
    g = g.download()
    loss_o, g_o = oracle.mlp_loss_grad(dims, params.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                                       n_real=45, emulate_bf16=True)
    assert abs(loss - loss_o) < 1e-4 * loss_o
    for l, ((gW, gb), (oW, ob)) in enumerate(zip(_split(g, dims), _split(g_o, dims))):
        print("layer", l, "rel err W", _rel(gW, oW), "b", _rel(gb, ob), flush=True)
        assert _rel(gW, oW) < 5e-3 and _rel(gb, ob) < 5e-3
    # padded rows of the head and padded columns of the input get no gradient
    (gW0, _), (gW4, gb4) = _split(g, dims)[0], _split(g, dims)[-1]
    assert not gW4[45:].any() and not gb4[45:].any()


def _linear_start(oracle, c, u):
    """(DS(u_0), DU(DS(u_0))^{-1} u_1, ...) for one cell (M, S)."""
    ds = oracle.DS(c, u[0])
    du = oracle.DU(c, ds)
    lam = np.zeros_like(u)
    lam[0] = ds
    lam[1:] = np.linalg.solve(du, u[1:].T).T
    return lam


def test_features_and_targets_follow_definition(oracle):
    b, cl = _closure_c2()
    net = native.WarmStartNet(cl, hidden=64, seed=1)
    u, lam_star, lam0 = workloads.closure_euler1d(512)
    du, dl = native.DeviceArray.from_host(u), native.DeviceArray.from_host(lam_star)
    x, t = net.features(du, dl, standardise=False)
    x, t = x.download(), t.download()
    c = oracle.closure(oracle.EULER1D, du_exact=1)
    lins = {}
    for j in (0, 17, 511):
        rho, m, E = u[j, 0]
        T = 0.4 * (E - 0.5 * m * m / rho) / rho
        D = np.array([rho, rho * np.sqrt(T), rho * T])
        np.testing.assert_allclose(x[j, :45], (u[j] / D).ravel(), rtol=2e-7)
        assert not x[j, 45:].any()
        lins[j] = _linear_start(oracle, c, u[j])
        np.testing.assert_allclose(t[j, :45], ((lam_star[j] - lins[j]) * D / rho).ravel(), rtol=2e-5, atol=1e-6)
    # the linearised start is one full Newton step from the plain start
    # (the oracle's Newton with alpha = 1 and one iteration)
    one = oracle.closure(oracle.EULER1D, du_exact=1, alpha=1.0, max_iter=1)
    l1 = lam0[17:18].copy()
    oracle.solve_closure(one, b.phi, b.phiw, u[17:18], l1)
    np.testing.assert_allclose(l1[0], lins[17], rtol=1e-9, atol=1e-12)
    # the inverse map: with a zero network (y = 0, mu_t = 0, sd_t = 1) the
    # prediction is the linearised start itself
    p, nrm = net.get_params()
    net.set_params(np.zeros_like(p), nrm)
    lp = net.predict(du).download()
    for j in lins:
        np.testing.assert_allclose(lp[j], lins[j], rtol=1e-12, atol=1e-14)


def test_select_keeps_the_better_start(oracle):
    b, cl = _closure_c2()
    u, lam_star, lam0 = workloads.closure_euler1d(1024)
    du = native.DeviceArray.from_host(u)
    lam = native.DeviceArray.from_host(lam0)
    alt = lam_star.copy()
    alt[::2] = lam0[::2] + 0.3  # worse on even cells, exact on odd cells
    alt[5] = np.nan             # non-finite never wins
    print("case: warm start select 1024 cells", flush=True)
    took = native.warm_start_select(cl, du, lam, native.DeviceArray.from_host(alt))
    out = lam.download()
    assert took == 511
    np.testing.assert_array_equal(out[1::2][[i for i in range(512) if 2 * i + 1 != 5]],
                                  lam_star[1::2][[i for i in range(512) if 2 * i + 1 != 5]])
    np.testing.assert_array_equal(out[::2], lam0[::2])
    np.testing.assert_array_equal(out[5], lam0[5])


def test_trained_warm_start_cuts_newton_steps():
    """Train on Newton-solved pairs of one seed, test on cells of another: the
    selected warm start must not change the solution and must save Newton steps.
    With the reference's damping alpha = 0.5 (Euler1D) Newton converges
    linearly at rate 1/2, so a start F times closer saves log2(F) steps: the
    gain is a few percent at this training size."""
    b, cl = _closure_c2()
    n_train, n_test = 1 << 16, 1 << 14
    u, lam_star, lam0 = workloads.closure_euler1d(n_train, seed=101, device=True)
    print("case: train warm start on", n_train, "pairs", flush=True)
    assert cl.solve(u, lam0) == 0  # the pairs come from the plain solver
    net = native.WarmStartNet(cl, hidden=256, seed=5)
    losses = []
    for lr in (1e-3, 5e-4, 2.5e-4, 1e-4):
        losses.append(net.train(u, lam0, batch=1024, steps=600, lr=lr, seed=len(losses) + 1))
        print("mean loss", losses[-1], flush=True)
    assert losses[-1] < 0.5 * losses[0]
    ut, ls_t, l0_t = workloads.closure_euler1d(n_test, seed=202, device=True)
    ref = native.DeviceArray.from_host(l0_t.download())
    assert cl.solve(ut, ref) == 0
    plain_steps, _ = cl.last_closure_stats()
    net.attach()
    warm = native.DeviceArray.from_host(l0_t.download())
    assert cl.solve(ut, warm) == 0
    warm_steps, _ = cl.last_closure_stats()
    net.detach()
    ls = ls_t.download()
    err_plain = np.abs(ref.download() - ls).max() / np.abs(ls).max()
    err_warm = np.abs(warm.download() - ls).max() / np.abs(ls).max()
    print("Newton steps plain", plain_steps, "warm", warm_steps, "lambda err", err_plain, err_warm, flush=True)
    assert err_warm < 1e-8 and err_plain < 1e-8
    assert warm_steps < 0.99 * plain_steps
