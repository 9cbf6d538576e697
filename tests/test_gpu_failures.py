"""Per-cell failure statuses of the GPU closure on moments that cannot be
closed -- unrealizable states (density <= 0), NaN and infinite moments --
mixed with ordinary cells, against the oracle's statuses (the reference
exits the process there: closure.cpp:180-183, :191-192, vectorspace.cpp:166-170;
SolveClosure accepts a NaN trial residual, closure.cpp:172, which ends here
with status 4).  Every kernel must finish with a status, never fault."""
import numpy as np
import pytest

from uqcreator_amd import native, workloads

pytestmark = pytest.mark.gpu


def _bad_cells(u):
    u = u.copy()
    u[0, 0, 0] = -0.5          # negative mean density: DS(u_0) is not finite
    u[1, 0, 0] = 0.0           # zero density
    u[2, 4, 1] = np.nan        # NaN in a higher moment: finite start, NaN gradient
    u[3, 0, 2] = np.nan        # NaN in the mean energy
    u[4, 2, 2] = np.inf        # infinite moment
    u[5, 0, 2] = -1.0          # negative energy: negative pressure
    return u, 6


@pytest.mark.parametrize("kernel,precision", [(native.KERNEL_GENERIC, native.FP64), (native.KERNEL_WAVE, native.FP64),
                                              (native.KERNEL_WAVE, native.MIXED), (native.KERNEL_BLOCK, native.FP64),
                                              (native.KERNEL_BLOCK, native.MIXED)])
@pytest.mark.parametrize("safe", [1, 0])
def test_failure_statuses_match_oracle(oracle, kernel, precision, safe):
    b = workloads.closure_basis()
    u, lam_star, lam0 = workloads.closure_euler1d(64, basis=b, seed=17)
    u, nbad = _bad_cells(u)
    c = oracle.closure(oracle.EULER1D, eps=1e-10, du_exact=1, safe=safe, max_iter=60)
    lam0 = lam0.copy()
    for j in range(nbad):
        lam0[j, 0] = oracle.DS(c, u[j, 0])  # SetDuals on the bad moments, NaN where DS is
    print("case: failures kernel", kernel, "precision", precision, "safe", safe, flush=True)
    cl = native.GpuClosure(native.EULER1D, b, epsilon=1e-10, du_exact=1, safe_line_search=safe, max_iterations=60,
                           precision=precision)
    cl.set_kernel(kernel)
    du = native.DeviceArray.from_host(u)
    dl = native.DeviceArray.from_host(lam0)
    st = native.DeviceArray(64, np.int32)
    it = native.DeviceArray(64, np.int32)
    nfail = cl.solve(du, dl, st, it)
    st, lam = st.download(), dl.download()
    assert cl.last_kernel() == kernel
    lam_o = lam0.copy()
    st_o, it_o = oracle.solve_closure(c, b.phi, b.phiw, u, lam_o)
    print("statuses GPU", st[:nbad].tolist(), "oracle", st_o[:nbad].tolist(), flush=True)
    np.testing.assert_array_equal(st, st_o)
    assert (st[:nbad] != 0).all()
    if safe:
        assert (st[nbad:] == 0).all()
    # (SolveClosure without the safe line search accepts a NaN trial in
    # some ordinary cells too; they end with status 4 on both sides)
    assert nfail == (st != 0).sum()
    ok = st == 0
    scale = np.abs(lam_star[ok]).max(axis=1, keepdims=True)
    assert (np.abs(lam[ok] - lam_o[ok]) / scale).max() < 1e-8


def test_bounded_barrier_outside_bounds(oracle):
    """Moments outside the barrier's bounds (u_0 beyond [u-, u+]) are not
    realizable: both the GPU and the oracle give a failure status for them
    and solve the other cells."""
    b = native.Basis(1, [5], 12)
    rng = np.random.default_rng(3)
    n = 64
    lam_star = rng.normal(0, 0.3, (n, b.M, 1))
    lam_star[:, 0] = rng.uniform(-2, 2, (n, 1))
    c = oracle.closure(oracle.BB, u_minus=2.9, u_plus=12.1, eps=1e-10, safe=1, max_iter=200)
    uq = oracle.eval_uq(c, b.phi, lam_star)
    u = oracle.reconstruct(1, b.phiw, np.ascontiguousarray(uq))
    u[0, 0, 0] = 13.0
    u[1, 0, 0] = 2.0
    u[2, 0, 0] = 12.1
    lam0 = np.zeros_like(u)
    lam_o = lam0.copy()
    st_o, _ = oracle.solve_closure(c, b.phi, b.phiw, u, lam_o)
    print("case: bounded barrier outside bounds", flush=True)
    for kernel in (native.KERNEL_GENERIC, native.KERNEL_WAVE):
        cl = native.GpuClosure(native.BOUNDED_BARRIER, b, u_minus=2.9, u_plus=12.1, epsilon=1e-10,
                               safe_line_search=1, max_iterations=200)
        cl.set_kernel(kernel)
        lam = lam0.copy()
        st, it = cl.solve_host(u, lam)
        print("kernel", kernel, "statuses GPU", st[:3].tolist(), "oracle", st_o[:3].tolist(), flush=True)
        assert (st[:3] != 0).all() and (st_o[:3] != 0).all()
        assert (st[3:] == 0).all() and (st_o[3:] == 0).all()
        scale = np.abs(lam_star[3:]).max(axis=1, keepdims=True)
        assert (np.abs(lam[3:] - lam_o[3:]) / scale).max() < 1e-8
