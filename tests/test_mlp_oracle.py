"""The FP64 oracle network that checks the device's warm-start kernels:
its gradient against central differences of its loss, GELU and GELU'
against their definitions, and the BF16 rounding against round-to-nearest-
even on known bit patterns."""
import math

import numpy as np


def test_gradient_matches_central_differences(oracle):
    dims = [8, 16, 16, 5]
    rng = np.random.default_rng(0)
    npar = oracle.mlp_param_count(dims)
    assert npar == 8 * 16 + 16 + 16 * 16 + 16 + 16 * 5 + 5
    p = rng.normal(0, 0.4, npar)
    x = rng.normal(size=(7, 8))
    t = rng.normal(size=(7, 5))
    loss, g = oracle.mlp_loss_grad(dims, p, x, t, n_real=4)
    h = 1e-6
    gn = np.zeros_like(g)
    for i in range(npar):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        gn[i] = (oracle.mlp_loss_grad(dims, pp, x, t, n_real=4)[0] - oracle.mlp_loss_grad(dims, pm, x, t, n_real=4)[0]) / (2 * h)
    assert np.abs(g - gn).max() < 1e-8 * max(1.0, np.abs(g).max())
    # the loss is the MSE of the forward pass over the first n_real outputs
    y = oracle.mlp_forward(dims, p, x)
    assert abs(((y[:, :4] - t[:, :4]) ** 2).mean() - loss) < 1e-14
    # the padded output gets no gradient: head row 4 and its bias
    off = 8 * 16 + 16 + 16 * 16 + 16
    assert not g[off + 4 * 16:off + 5 * 16].any() and g[off + 80 + 4] == 0.0


def test_gelu_tanh_form(oracle):
    for z in (-6.0, -1.3, -0.2, 0.0, 0.7, 2.5, 9.0):
        ref = 0.5 * z * (1 + math.tanh(math.sqrt(2 / math.pi) * (z + 0.044715 * z ** 3)))
        assert abs(oracle.gelu(z) - ref) < 1e-15 * max(1, abs(z))
        d = (oracle.gelu(z + 1e-6) - oracle.gelu(z - 1e-6)) / 2e-6
        assert abs(oracle.gelu(z, derivative=True) - d) < 1e-8
    # tends to the identity for large z and to 0 for very negative z
    assert abs(oracle.gelu(9.0) - 9.0) < 1e-12 and abs(oracle.gelu(-9.0)) < 1e-12


def test_bf16_round_to_nearest_even(oracle):
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even (1.0)
    assert oracle.bf16_round(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even (1 + 2^-6)
    assert oracle.bf16_round(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert oracle.bf16_round(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert oracle.bf16_round(-3.0) == -3.0
    v = np.random.default_rng(1).normal(size=1000)
    r = oracle.bf16_round(v)
    assert (np.abs(r - v) <= np.abs(v) * 2.0 ** -8 * 1.0000001).all()


def test_bf16_emulation_changes_results_by_bf16_error(oracle):
    dims = [32, 64, 32]
    rng = np.random.default_rng(2)
    p = rng.normal(0, 0.2, oracle.mlp_param_count(dims))
    x = rng.normal(size=(16, 32))
    y = oracle.mlp_forward(dims, p, x)
    yb = oracle.mlp_forward(dims, p, x, emulate_bf16=True)
    rel = np.linalg.norm(y - yb) / np.linalg.norm(y)
    assert 1e-4 < rel < 3e-2
