"""The allocation-free host closure entry point (uqk_closure_solve_host_noalloc)
against uqk_closure_solve_host, through the C ABI: the same kernels on the
same inputs, so duals, statuses and Newton counts must be identical.  Each
case calls the allocation-free entry point twice in a row, the second time
with fewer cells, so contents the first call left in the shared staging array
must not reach the second call's results."""
import numpy as np
import pytest

from uqcreator_amd import native, workloads

pytestmark = pytest.mark.gpu


def _euler2d_orders_case(oracle, n):
    """Euler 2D cells with per-cell orders 1..4 (nTotal 3..15, workgroup or
    generic kernel by size), moments from known duals with the oracle."""
    rng = np.random.default_rng(41)
    b = native.Basis(2, [1, 2, 3, 4], 7)
    c = oracle.closure(oracle.EULER2D, eps=1e-9, safe=1)
    ntl = np.asarray(b.ntotal_per_level)
    nt = ntl[rng.integers(0, len(ntl), n)].astype(np.int32)
    lam_star = np.zeros((n, b.M, 4))
    for j in range(n):
        rho, vx, vy, p = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2)
        lam_star[j, 0] = oracle.DS(c, np.array([rho, rho * vx, rho * vy, p / 0.4 + 0.5 * rho * (vx * vx + vy * vy)]))
        lam_star[j, 1:nt[j]] = rng.normal(0, 0.01, (nt[j] - 1, 4)) * np.abs(lam_star[j, 0])
    uq = oracle.eval_uq(c, b.phi, lam_star, nt_cell=nt)
    u = oracle.reconstruct(4, b.phiw, np.ascontiguousarray(uq), nt_cell=nt)
    lam0 = np.zeros_like(lam_star)
    for j in range(n):
        lam0[j, 0] = oracle.DS(c, u[j, 0])
    cl = native.GpuClosure(native.EULER2D, b, epsilon=1e-9, safe_line_search=1)
    return cl, u, lam0, lam_star, nt


@pytest.mark.parametrize("case", ["euler1d_c2_basis", "euler2d_per_cell_orders"])
def test_noalloc_host_solve_equals_host_solve(oracle, case):
    sizes = (3000, 700)
    if case == "euler1d_c2_basis":
        b = workloads.closure_basis()
        u, lam_star, lam0 = workloads.closure_euler1d(sizes[0], basis=b, seed=23)
        cl = native.GpuClosure(native.EULER1D, b, epsilon=1e-10, du_exact=1, safe_line_search=1)
        nt = None
    else:
        cl, u, lam0, lam_star, nt = _euler2d_orders_case(oracle, sizes[0])
    print("case: no-alloc host solve", case, "sizes", sizes, flush=True)
    got = []
    for n in sizes:  # two calls in a row, the second smaller
        lam = np.ascontiguousarray(lam0[:n]).copy()
        st, it = cl.solve_host_noalloc(np.ascontiguousarray(u[:n]), lam, nt_cell=None if nt is None else nt[:n])
        got.append((lam, st, it))
    for n, (lam, st, it) in zip(sizes, got):
        lam_ref = np.ascontiguousarray(lam0[:n]).copy()
        st_ref, it_ref = cl.solve_host(np.ascontiguousarray(u[:n]), lam_ref, nt_cell=None if nt is None else nt[:n])
        assert (st_ref == 0).all()
        np.testing.assert_array_equal(st, st_ref)
        np.testing.assert_array_equal(it, it_ref)
        np.testing.assert_array_equal(lam, lam_ref)
        scale = np.abs(lam_star[:n]).max(axis=1, keepdims=True)
        err = (np.abs(lam - lam_star[:n]) / scale).max()
        print(case, n, "cells: max rel |lambda - lambda*| = %.2e" % err, flush=True)
        assert err < 1e-7
        if nt is not None:
            for j in range(n):
                assert np.all(lam[j, nt[j]:] == 0.0)


def test_noalloc_rejects_what_it_cannot_stage():
    """Batches beyond the fixed staging array and contexts with a learned
    warm start attached are refused with an error, before any GPU work."""
    b = workloads.closure_basis()
    cl = native.GpuClosure(native.EULER1D, b, epsilon=1e-10, du_exact=1, safe_line_search=1)
    n = 100000  # 100,000 cells x 2 x 45 duals x 8 B = 72 MB > 64 MiB
    u = np.zeros((n, b.M, 3))
    lam = np.zeros((n, b.M, 3))
    with pytest.raises(RuntimeError, match="MiB"):
        cl.solve_host_noalloc(u, lam)
    net = native.WarmStartNet(cl, hidden=64, seed=1)
    net.attach()
    u1, _, lam1 = workloads.closure_euler1d(64, basis=b, seed=29)
    with pytest.raises(RuntimeError, match="warm start"):
        cl.solve_host_noalloc(u1, lam1.copy())
    net.detach()
    st, _ = cl.solve_host_noalloc(u1, lam1.copy())
    assert (st == 0).all()
