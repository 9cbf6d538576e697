"""ctypes wrapper of the FP64 CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always to check results against or as the timed CPU baseline.
The product package (uqcreator_amd) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

BB, EULER1D, EULER2D, SW2D, SW2D_BARRIER = range(5)
P_BURGERS, P_EULER2D, P_SW2D = range(3)
FLUX_REF, FLUX_LF = range(2)
BC_NONE, BC_DIRICHLET, BC_NOSLIP, BC_SWWALL = range(4)


class Closure(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("S", C.c_int), ("gamma", C.c_double), ("u_minus", C.c_double),
        ("u_plus", C.c_double), ("g", C.c_double), ("barrier", C.c_double), ("alpha", C.c_double),
        ("eps", C.c_double), ("max_iter", C.c_int), ("max_halvings", C.c_int), ("du_exact", C.c_int),
        ("safe", C.c_int),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("problem", C.c_int), ("flux", C.c_int), ("cfl", C.c_double), ("gamma", C.c_double),
        ("g", C.c_double), ("friction", C.c_double), ("friction_dim", C.c_int),
    ]


class Mesh(C.Structure):
    _fields_ = [
        ("ncell", C.c_long), ("nedge", C.c_int), ("dim", C.c_int), ("nbr", C.c_void_p),
        ("unit_normal", C.c_void_p), ("normal", C.c_void_p), ("area", C.c_void_p),
        ("dx_param", C.c_void_p), ("bc", C.c_void_p), ("bnormal", C.c_void_p),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        L = _lib
        L.orc_legendre.restype = C.c_double
        L.orc_legendre.argtypes = [C.c_int, C.c_double]
        L.orc_mlp_param_count.restype = C.c_long
        L.orc_gelu.restype = C.c_double
        L.orc_gelu.argtypes = [C.c_double, C.c_int]
        L.orc_bf16_round.restype = C.c_double
        L.orc_bf16_round.argtypes = [C.c_double]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_compute_dt.restype = C.c_double
        L.orc_compute_dt.argtypes = [C.POINTER(Problem), C.c_int, C.c_int, C.c_void_p, C.c_double]
        L.orc_ref_indicator.restype = C.c_double
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def closure(kind, S=None, gamma=1.4, u_minus=3.0, u_plus=12.0, g=9.81, barrier=0.0, alpha=None,
            eps=1e-10, max_iter=1000, max_halvings=1000, du_exact=0, safe=0):
    default_S = {BB: 1, EULER1D: 3, EULER2D: 4, SW2D: 3, SW2D_BARRIER: 3}[kind]
    default_alpha = {BB: 1.0, EULER1D: 0.5, EULER2D: 1.0, SW2D: 1.0, SW2D_BARRIER: 1.0}[kind]
    return Closure(kind, S or default_S, gamma, u_minus, u_plus, g, barrier,
                   default_alpha if alpha is None else alpha, eps, max_iter, max_halvings, du_exact, safe)


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _p(x), _p(w))
    return x, w


def legendre(n, x):
    return lib().orc_legendre(n, float(x))


def index_set(d, degrees, max_degree=False):
    degrees = np.asarray(degrees, dtype=np.int32)
    cap = (int(degrees[-1]) + 1) ** d
    idx = np.zeros(cap * d, dtype=np.int32)
    ntl = np.zeros(len(degrees), dtype=np.int32)
    n = lib().orc_index_set(d, len(degrees), _p(degrees), int(max_degree), cap, _p(idx), _p(ntl))
    return idx[: n * d].reshape(n, d), ntl


def tables(d, q1, idx):
    M = idx.shape[0]
    Q = q1 ** d
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    xi = np.zeros((Q, d))
    wq = np.zeros(Q)
    phi = np.zeros((Q, M))
    phiw = np.zeros((Q, M))
    lib().orc_tables(d, q1, M, _p(idx), _p(xi), _p(wq), _p(phi), _p(phiw))
    return xi, wq, phi, phiw


def U(c, lam):
    out = np.zeros(c.S)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    lib().orc_U(C.byref(c), _p(lam), _p(out))
    return out


def U_mat(c, lam):
    out = np.zeros(c.S)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    lib().orc_U_mat(C.byref(c), _p(lam), _p(out))
    return out


def DU(c, lam):
    y = np.zeros(c.S * c.S)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    lib().orc_DU(C.byref(c), _p(lam), _p(y))
    return y.reshape(c.S, c.S)


def DS(c, u):
    out = np.zeros(c.S)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lib().orc_DS(C.byref(c), _p(u), _p(out))
    return out


def gradient(c, phi, phiw, lam, u, nt=None):
    Q, M = phi.shape
    nt = M if nt is None else nt
    g = np.zeros(c.S * nt)
    lib().orc_gradient(C.byref(c), M, Q, nt, _p(phi), _p(phiw), _p(np.ascontiguousarray(lam)),
                       _p(np.ascontiguousarray(u)), _p(g))
    return g


def hessian(c, phi, phiw, lam, nt=None):
    Q, M = phi.shape
    nt = M if nt is None else nt
    n = c.S * nt
    H = np.zeros(n * n)
    lib().orc_hessian(C.byref(c), M, Q, nt, _p(phi), _p(phiw), _p(np.ascontiguousarray(lam)), _p(H))
    return H.reshape(n, n).T.copy()  # column-major -> (row, col)


def posv_upper(H, b):
    """Returns (info, x, U) with H given as a (row, col) array."""
    n = H.shape[0]
    A = np.ascontiguousarray(H.T).copy()  # column-major storage
    x = np.array(b, dtype=np.float64).copy()
    info = lib().orc_posv_upper(n, _p(A), _p(x))
    Umat = np.triu(A.reshape(n, n).T)
    return info, x, Umat


def solve_closure(c, phi, phiw, u, lam, nt_cell=None, nthreads=0):
    """u, lam: (ncell, M, S) arrays (reference layout).  lam is updated in place."""
    Q, M = phi.shape
    ncell = u.shape[0]
    assert u.flags.c_contiguous and lam.flags.c_contiguous
    status = np.zeros(ncell, dtype=np.int32)
    iters = np.zeros(ncell, dtype=np.int32)
    ntc = None if nt_cell is None else np.ascontiguousarray(nt_cell, dtype=np.int32)
    lib().orc_solve_closure(C.byref(c), M, Q, _p(phi), _p(phiw), C.c_long(ncell), _p(ntc), _p(u), _p(lam),
                            _p(status), _p(iters), nthreads)
    return status, iters


def eval_uq(c, phi, lam, nt_cell=None, nthreads=0):
    Q, M = phi.shape
    ncell = lam.shape[0]
    uq = np.zeros((ncell, Q, c.S))
    ntc = None if nt_cell is None else np.ascontiguousarray(nt_cell, dtype=np.int32)
    lib().orc_eval_uq(C.byref(c), M, Q, _p(phi), C.c_long(ncell), _p(ntc), _p(np.ascontiguousarray(lam)), _p(uq),
                      nthreads)
    return uq


def reconstruct(S, phiw, uq, nt_cell=None, nthreads=0):
    Q, M = phiw.shape
    ncell = uq.shape[0]
    u = np.zeros((ncell, M, S))
    ntc = None if nt_cell is None else np.ascontiguousarray(nt_cell, dtype=np.int32)
    lib().orc_reconstruct(S, M, Q, _p(phiw), C.c_long(ncell), _p(ntc), _p(np.ascontiguousarray(uq)), _p(u), nthreads)
    return u


def mean_var(c, phi, phiw, lam, nt_cell=None, nthreads=0):
    Q, M = phi.shape
    ncell = lam.shape[0]
    mv = np.zeros((ncell, 2 * c.S))
    ntc = None if nt_cell is None else np.ascontiguousarray(nt_cell, dtype=np.int32)
    lib().orc_mean_var(C.byref(c), M, Q, _p(phi), _p(phiw), C.c_long(ncell), _p(ntc), _p(np.ascontiguousarray(lam)),
                       _p(mv), nthreads)
    return mv


def problem(kind, flux=FLUX_REF, cfl=0.5, gamma=1.4, g=9.81, friction=0.0, friction_dim=-1):
    return Problem(kind, flux, cfl, gamma, g, friction, friction_dim)


def flux_point(p, u, v, nu, n):
    S = len(u)
    out = np.zeros(S)
    dim = len(nu)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, nu, n)]
    lib().orc_flux_point(C.byref(p), S, *[_p(a) for a in arrs], dim, _p(out))
    return out


def compute_dt(p, uq, dx):
    Q, S = uq.shape
    uq = np.ascontiguousarray(uq, dtype=np.float64)
    return lib().orc_compute_dt(C.byref(p), S, Q, _p(uq), float(dx))


class MeshArrays:
    """Holds the numpy arrays behind an orc_mesh struct."""

    def __init__(self, nbr, unit_normal, normal, area, dx_param, bc, bnormal):
        self.nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.unit_normal = np.ascontiguousarray(unit_normal, dtype=np.float64)
        self.normal = np.ascontiguousarray(normal, dtype=np.float64)
        self.area = np.ascontiguousarray(area, dtype=np.float64)
        self.dx_param = np.ascontiguousarray(dx_param, dtype=np.float64)
        self.bc = np.ascontiguousarray(bc, dtype=np.int32)
        self.bnormal = np.ascontiguousarray(bnormal, dtype=np.float64)
        ncell, nedge = self.nbr.shape
        dim = self.unit_normal.shape[-1]
        self.struct = Mesh(ncell, nedge, dim, _p(self.nbr), _p(self.unit_normal), _p(self.normal), _p(self.area),
                           _p(self.dx_param), _p(self.bc), _p(self.bnormal))


def advance(p, mesh, phiw, xi, dt, u, uq, nt_cell=None, nthreads=0):
    Q, M = phiw.shape
    S = u.shape[2]
    d = xi.shape[1]
    unew = np.zeros_like(u)
    ntc = None if nt_cell is None else np.ascontiguousarray(nt_cell, dtype=np.int32)
    lib().orc_advance(C.byref(p), C.byref(mesh.struct), S, M, Q, d, _p(phiw), _p(np.ascontiguousarray(xi)), _p(ntc),
                      C.c_double(dt), _p(np.ascontiguousarray(u)), _p(np.ascontiguousarray(uq)), _p(unew), nthreads)
    return unew


def run(c, p, mesh, phi, phiw, xi, u, lam, t_end, max_steps=-1, nt_cell=None, nthreads=0):
    """Time loop; u and lam (ncell, M, S) are updated in place."""
    Q, M = phi.shape
    d = xi.shape[1]
    t = C.c_double()
    steps = C.c_long()
    res = C.c_double()
    newton = C.c_long()
    ntc = None if nt_cell is None else np.ascontiguousarray(nt_cell, dtype=np.int32)
    rc = lib().orc_run(C.byref(c), C.byref(p), C.byref(mesh.struct), M, Q, d, _p(phi), _p(phiw),
                       _p(np.ascontiguousarray(xi)), _p(ntc), _p(u), _p(lam), C.c_double(t_end), C.c_long(max_steps),
                       C.byref(t), C.byref(steps), C.byref(res), C.byref(newton), nthreads)
    return dict(rc=rc, t=t.value, steps=steps.value, residual=res.value, newton_iters=newton.value)


def ref_indicator(S, d, nt_level, lev, u_cell):
    """momentsolver.cpp:271-289 for one cell's moments u_cell (M, S)."""
    lv = np.ascontiguousarray(nt_level, dtype=np.int32)
    return lib().orc_ref_indicator(S, d, _p(lv), int(lev), _p(np.ascontiguousarray(u_cell, dtype=np.float64)))


def refine(u, d, nt_level, nt_cell, above=0.02, below=0.004, max_level=None):
    """momentsolver.cpp:104-117 over all cells; returns the new nt_cell."""
    ncell, M, S = u.shape
    lv = np.ascontiguousarray(nt_level, dtype=np.int32)
    nt = np.array(nt_cell, dtype=np.int32, copy=True)
    cap = len(lv) - 1 if max_level is None else int(max_level)
    lib().orc_refine(S, M, d, len(lv), _p(lv), C.c_double(above), C.c_double(below), cap, C.c_long(ncell),
                     _p(np.ascontiguousarray(u, dtype=np.float64)), _p(nt))
    return nt


def run_levels(c, p, mesh, phi, phiw, xi, u, lam, nt_level, nt_cell, t_end, max_steps=-1, above=0.02, below=0.004,
               max_level=None, nthreads=0):
    """Time loop with refinement levels; u, lam (ncell, M, S) and nt_cell
    (int32 array, updated) in place."""
    Q, M = phi.shape
    d = xi.shape[1]
    t = C.c_double()
    steps = C.c_long()
    res = C.c_double()
    newton = C.c_long()
    lv = np.ascontiguousarray(nt_level, dtype=np.int32)
    cap = len(lv) - 1 if max_level is None else int(max_level)
    assert nt_cell.dtype == np.int32 and nt_cell.flags.c_contiguous
    rc = lib().orc_run_levels(C.byref(c), C.byref(p), C.byref(mesh.struct), M, Q, d, _p(phi), _p(phiw),
                              _p(np.ascontiguousarray(xi)), len(lv), _p(lv), C.c_double(above), C.c_double(below),
                              cap, _p(nt_cell), _p(u), _p(lam), C.c_double(t_end), C.c_long(max_steps), C.byref(t),
                              C.byref(steps), C.byref(res), C.byref(newton), nthreads)
    return dict(rc=rc, t=t.value, steps=steps.value, residual=res.value, newton_iters=newton.value)


def rel_error(sol, ref, area, lnorm=2, mask=None):
    ncell, nf = sol.shape
    err = np.zeros(nf)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_rel_error(nf, C.c_long(ncell), _p(np.ascontiguousarray(sol)), _p(np.ascontiguousarray(ref)),
                        _p(np.ascontiguousarray(area)), _p(m), lnorm, _p(err))
    return err


def _mlp_dims(dims):
    return np.ascontiguousarray(dims, dtype=np.int32)


def mlp_param_count(dims):
    d = _mlp_dims(dims)
    return lib().orc_mlp_param_count(len(d) - 1, _p(d))


def gelu(z, derivative=False):
    return np.vectorize(lambda v: lib().orc_gelu(C.c_double(v), int(derivative)))(z)


def bf16_round(v):
    return np.vectorize(lambda a: lib().orc_bf16_round(C.c_double(a)))(v)


def mlp_forward(dims, params, x, emulate_bf16=False):
    d = _mlp_dims(dims)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros((x.shape[0], d[-1]))
    lib().orc_mlp_forward(len(d) - 1, _p(d), _p(np.ascontiguousarray(params, dtype=np.float64)), C.c_long(x.shape[0]),
                          _p(x), _p(y), int(emulate_bf16))
    return y


def mlp_loss_grad(dims, params, x, t, n_real=None, emulate_bf16=False):
    d = _mlp_dims(dims)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.zeros(mlp_param_count(dims))
    loss = lib().orc_mlp_loss_grad(len(d) - 1, _p(d), _p(np.ascontiguousarray(params, dtype=np.float64)),
                                   C.c_long(x.shape[0]), _p(x), _p(t), int(d[-1] if n_real is None else n_real),
                                   _p(g), int(emulate_bf16))
    return loss, g
