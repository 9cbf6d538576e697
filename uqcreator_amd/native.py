"""ctypes binding of the C ABI in include/uqk.h (libuqk.so).

The library must have been built (``python -m uqcreator_amd.build`` or
``__graft_entry__.build()``); loading fails loudly otherwise -- there is no
CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libuqk.so")

BOUNDED_BARRIER, EULER1D, EULER2D, SHALLOW_WATER2D, SHALLOW_WATER2D_BARRIER = range(5)
FP64, MIXED = 0, 1
BURGERS, EULER2D_PROBLEM, SHALLOW_WATER2D_PROBLEM = range(3)
FLUX_REFERENCE, FLUX_LAX_FRIEDRICHS = 0, 1
BC_NONE, BC_DIRICHLET, BC_NOSLIP, BC_SWWALL = range(4)
KERNEL_AUTO, KERNEL_GENERIC, KERNEL_WAVE, KERNEL_BLOCK = range(4)
STATES = {BOUNDED_BARRIER: 1, EULER1D: 3, EULER2D: 4, SHALLOW_WATER2D: 3, SHALLOW_WATER2D_BARRIER: 3}
# Newton damping per closure (boundedbarrier.cpp:4, eulerclosure.cpp:3,
# eulerclosure2d.cpp:3, shallowwaterclosure2d.cpp:3-6)
ALPHA = {BOUNDED_BARRIER: 1.0, EULER1D: 0.5, EULER2D: 1.0, SHALLOW_WATER2D: 1.0, SHALLOW_WATER2D_BARRIER: 1.0}


class Desc(C.Structure):
    _fields_ = [
        ("closure", C.c_int32), ("n_states", C.c_int32), ("n_moments", C.c_int32), ("n_quad", C.c_int32),
        ("dim_xi", C.c_int32), ("gamma", C.c_double), ("u_minus", C.c_double), ("u_plus", C.c_double),
        ("gravity", C.c_double), ("barrier", C.c_double), ("alpha", C.c_double), ("epsilon", C.c_double),
        ("max_iterations", C.c_int32), ("max_halvings", C.c_int32), ("du_exact", C.c_int32),
        ("safe_line_search", C.c_int32), ("precision", C.c_int32), ("phi", C.c_void_p), ("phiw", C.c_void_p),
        ("xi", C.c_void_p), ("device", C.c_int32),
    ]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("problem", C.c_int32), ("flux", C.c_int32), ("cfl", C.c_double), ("gamma", C.c_double),
        ("gravity", C.c_double), ("friction", C.c_double), ("friction_dim", C.c_int32),
    ]


class RunReport(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int64), ("steps", C.c_int64), ("newton_steps", C.c_int64), ("t", C.c_double),
        ("residual", C.c_double), ("runtime_s", C.c_double), ("cell_steps_per_s", C.c_double),
        ("n_states", C.c_int32), ("n_moments", C.c_int32), ("n_quad", C.c_int32),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int64), ("n_edges", C.c_int32), ("dim", C.c_int32), ("neighbor", C.c_void_p),
        ("unit_normal", C.c_void_p), ("normal", C.c_void_p), ("area", C.c_void_p), ("dx_param", C.c_void_p),
        ("bc", C.c_void_p), ("boundary_normal", C.c_void_p),
    ]


# Every symbol declared in include/uqk.h, with its ctypes signature.
SIGNATURES = {
    "uqk_create": (C.c_int, [C.POINTER(Desc), C.POINTER(C.c_void_p)]),
    "uqk_destroy": (C.c_int, [C.c_void_p]),
    "uqk_last_error": (C.c_char_p, [C.c_void_p]),
    "uqk_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "uqk_synchronize": (C.c_int, [C.c_void_p]),
    "uqk_set_closure_kernel": (C.c_int, [C.c_void_p, C.c_int32]),
    "uqk_last_closure_kernel": (C.c_int32, [C.c_void_p]),
    "uqk_closure_solve": (C.c_int64, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "uqk_closure_solve_host": (C.c_int64, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "uqk_closure_solve_host_noalloc": (C.c_int64, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                                   C.c_void_p, C.c_void_p]),
    "uqk_eval_uq": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_reconstruct": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_compute_dt": (C.c_int, [C.c_void_p, C.POINTER(ProblemDesc), C.POINTER(MeshDesc), C.c_void_p,
                                 C.POINTER(C.c_double)]),
    "uqk_advance": (C.c_int, [C.c_void_p, C.POINTER(ProblemDesc), C.POINTER(MeshDesc), C.c_void_p, C.c_double,
                              C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_step": (C.c_int64, [C.c_void_p, C.POINTER(ProblemDesc), C.POINTER(MeshDesc), C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                             C.POINTER(C.c_double)]),
    "uqk_mean_var": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_set_refinement": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_double, C.c_double]),
    "uqk_refine": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "uqk_set_refinement_cap": (C.c_int, [C.c_void_p, C.c_int32]),
    "uqk_run_config_output": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(RunReport), C.c_char_p,
                                        C.c_int32]),
    "uqk_su2_mesh": (C.c_int64, [C.c_char_p, C.c_char_p, C.POINTER(C.c_int32), C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p, C.c_int32]),
    "uqk_run_steps": (C.c_int64, [C.c_void_p, C.POINTER(ProblemDesc), C.POINTER(MeshDesc), C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                  C.c_int64, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "uqk_gauss_legendre": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p]),
    "uqk_index_set": (C.c_int32, [C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    "uqk_basis_tables": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "uqk_last_closure_ms": (C.c_double, [C.c_void_p]),
    "uqk_last_closure_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "uqk_set_kernel_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "uqk_set_factor_reuse": (C.c_int, [C.c_void_p, C.c_int32]),
    "uqk_set_residual_check": (C.c_int, [C.c_void_p, C.c_int32]),
    "uqk_invalidate_factors": (C.c_int, [C.c_void_p]),
    "uqk_last_closure_factorizations": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "uqk_kernel_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "uqk_synth_euler1d_closure": (C.c_int64, [C.c_int64, C.c_uint64, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_synth_euler1d_closure_host": (C.c_int64, [C.c_int64, C.c_uint64, C.c_int32, C.c_int32, C.c_void_p,
                                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_device_alloc": (C.c_void_p, [C.c_int64]),
    "uqk_device_free": (C.c_int, [C.c_void_p]),
    "uqk_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32]),
    "uqk_device_synchronize": (C.c_int, []),
    "uqk_device_count": (C.c_int32, []),
    "uqk_set_launch_shape": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "uqk_get_launch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "uqk_autotune": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_char_p]),
    "uqk_load_tuning": (C.c_int, [C.c_void_p, C.c_char_p]),
    "uqk_mlp_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "uqk_mlp_destroy": (C.c_int, [C.c_void_p]),
    "uqk_mlp_shape": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]),
    "uqk_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_mlp_fit_normalization": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "uqk_mlp_features": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                   C.c_void_p]),
    "uqk_mlp_forward": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "uqk_mlp_loss_grad": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32,
                                    C.POINTER(C.c_double), C.c_void_p]),
    "uqk_mlp_train": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                C.c_uint64, C.POINTER(C.c_double)]),
    "uqk_mlp_predict": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "uqk_warm_start_select": (C.c_int64, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "uqk_set_warm_start": (C.c_int, [C.c_void_p, C.c_void_p]),
    "uqk_run_config": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(RunReport), C.c_char_p, C.c_int32]),
    "uqk_config_setup": (C.c_int64, [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_char_p,
                                     C.c_int32]),
}

_lib = None


def lib():
    """Load libuqk.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libuqk.so is not built (%s); run python -m uqcreator_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    if isinstance(a, DeviceArray):
        return C.c_void_p(a.ptr)
    return C.c_void_p(a.data_ptr())  # any object exposing a device pointer


class DeviceArray:
    """A device buffer owned through the C ABI (uqk_device_alloc), with the
    shape/dtype of the host array it mirrors."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self.ptr = lib().uqk_device_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError("uqk_device_alloc(%d) failed" % self.nbytes)

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if lib().uqk_memcpy(C.c_void_p(self.ptr), _p(a), self.nbytes, 0) != 0:
            raise RuntimeError("host->device copy failed")

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if lib().uqk_memcpy(_p(out), C.c_void_p(self.ptr), self.nbytes, 1) != 0:
            raise RuntimeError("device->host copy failed")
        return out

    def copy_from(self, other, dst_offset=None):
        """Device copy of `other`: the whole array, or (dst_offset given) into
        the rows dst_offset.. of the first dimension."""
        if dst_offset is None:
            assert other.nbytes == self.nbytes
            dst = self.ptr
        else:
            row = self.nbytes // self.shape[0]
            assert other.nbytes % row == 0 and 0 <= dst_offset and dst_offset * row + other.nbytes <= self.nbytes
            dst = self.ptr + dst_offset * row
        if lib().uqk_memcpy(C.c_void_p(dst), C.c_void_p(other.ptr), other.nbytes, 2) != 0:
            raise RuntimeError("device->device copy failed")

    def free(self):
        if getattr(self, "ptr", None):
            lib().uqk_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count():
    return lib().uqk_device_count()


def synchronize():
    if lib().uqk_device_synchronize() != 0:
        raise RuntimeError("device synchronize failed")


# ---------------------------------------------------------------------------
# host numerics (Settings index set, quadrature, Closure tables)
# ---------------------------------------------------------------------------
def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    if lib().uqk_gauss_legendre(n, _p(x), _p(w)) != 0:
        raise RuntimeError("gauss_legendre failed")
    return x, w


def index_set(d, degrees, max_degree=False):
    deg = np.asarray(degrees, dtype=np.int32)
    cap = (int(deg[-1]) + 1) ** d
    idx = np.zeros(cap * d, dtype=np.int32)
    ntl = np.zeros(len(deg), dtype=np.int32)
    n = lib().uqk_index_set(d, len(deg), _p(deg), int(max_degree), cap, _p(idx), _p(ntl))
    if n < 0:
        raise RuntimeError("index_set failed")
    return idx[: n * d].reshape(n, d), ntl


def basis_tables(d, q1, idx):
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    M = idx.shape[0]
    Q = q1 ** d
    xi = np.zeros((Q, d))
    wq = np.zeros(Q)
    phi = np.zeros((Q, M))
    phiw = np.zeros((Q, M))
    r = lib().uqk_basis_tables(d, q1, M, _p(idx), _p(xi), _p(wq), _p(phi), _p(phiw))
    if r != Q:
        raise RuntimeError("basis_tables failed")
    return xi, wq, phi, phiw


class Basis:
    """Index set, quadrature and the Closure's phi / phiw tables."""

    def __init__(self, d, degrees, q1, max_degree=False):
        self.d = d
        self.degrees = list(degrees)
        self.q1 = q1
        self.idx, self.ntotal_per_level = index_set(d, degrees, max_degree)
        self.xi, self.wq, self.phi, self.phiw = basis_tables(d, q1, self.idx)
        self.Q, self.M = self.phi.shape


# ---------------------------------------------------------------------------
# GPU closure context
# ---------------------------------------------------------------------------
class GpuClosure:
    """The GPU Closure: uqk_create / uqk_closure_solve / sweeps, one context."""

    def __init__(self, kind, basis, gamma=1.4, u_minus=3.0, u_plus=12.0, gravity=9.81, barrier=0.0, alpha=None,
                 epsilon=1e-10, max_iterations=1000, max_halvings=1000, du_exact=1, safe_line_search=0,
                 precision=FP64, device=0):
        self.kind = kind
        self.S = STATES[kind]
        self.basis = basis
        self.M, self.Q = basis.M, basis.Q
        self._phi = np.ascontiguousarray(basis.phi)
        self._phiw = np.ascontiguousarray(basis.phiw)
        self._xi = np.ascontiguousarray(basis.xi)
        self.desc = Desc(kind, self.S, self.M, self.Q, basis.d, gamma, u_minus, u_plus, gravity, barrier,
                         ALPHA[kind] if alpha is None else alpha, epsilon, max_iterations, max_halvings, du_exact,
                         safe_line_search, precision, _p(self._phi), _p(self._phiw), _p(self._xi), device)
        h = C.c_void_p()
        rc = lib().uqk_create(C.byref(self.desc), C.byref(h))
        self.h = h
        if rc != 0:
            msg = lib().uqk_last_error(h).decode() if h else "uqk_create failed"
            self.close()
            raise RuntimeError("uqk_create: %s (%d)" % (msg, rc))

    def close(self):
        if getattr(self, "h", None):
            lib().uqk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc < 0:
            raise RuntimeError("%s: %s (%d)" % (what, lib().uqk_last_error(self.h).decode(), rc))
        return rc

    def error(self):
        return lib().uqk_last_error(self.h).decode()

    def set_launch_shape(self, waves=0, blocks_per_cu=0):
        self._check(lib().uqk_set_launch_shape(self.h, waves, blocks_per_cu), "uqk_set_launch_shape")

    def launch_shape(self):
        w, b = C.c_int32(), C.c_int32()
        lib().uqk_get_launch_shape(self.h, C.byref(w), C.byref(b))
        return w.value, b.value

    def autotune(self, u, lam, json_path=None):
        """Times every launch shape of the specialised kernel on these cells
        and keeps the fastest; returns the number of shapes timed."""
        return self._check(lib().uqk_autotune(self.h, u.shape[0], _p(u), _p(lam),
                                              json_path.encode() if json_path else None), "uqk_autotune")

    def load_tuning(self, json_path):
        return self._check(lib().uqk_load_tuning(self.h, json_path.encode()), "uqk_load_tuning")

    def set_kernel(self, kernel):
        """KERNEL_AUTO, KERNEL_GENERIC, KERNEL_WAVE or KERNEL_BLOCK."""
        self._check(lib().uqk_set_closure_kernel(self.h, kernel), "uqk_set_closure_kernel")

    def last_kernel(self):
        return lib().uqk_last_closure_kernel(self.h)

    def set_stream(self, stream_handle):
        self._check(lib().uqk_set_stream(self.h, C.c_void_p(stream_handle)), "uqk_set_stream")

    def solve(self, u, lam, status=None, iters=None, nt_cell=None):
        """Device tensors (ncell, M, S) float64; lam updated in place. Returns #failed cells."""
        n = u.shape[0]
        return self._check(lib().uqk_closure_solve(self.h, n, _p(nt_cell), _p(u), _p(lam), _p(status), _p(iters)),
                           "uqk_closure_solve")

    def solve_host(self, u, lam, nt_cell=None):
        """numpy (ncell, M, S) float64; lam updated in place. Returns (status, iters)."""
        n = u.shape[0]
        assert u.flags.c_contiguous and lam.flags.c_contiguous
        status = np.zeros(n, dtype=np.int32)
        iters = np.zeros(n, dtype=np.int32)
        ntc = None if nt_cell is None else np.ascontiguousarray(nt_cell, dtype=np.int32)
        self._check(lib().uqk_closure_solve_host(self.h, n, _p(ntc), _p(u), _p(lam), _p(status), _p(iters)),
                    "uqk_closure_solve_host")
        return status, iters

    def solve_host_noalloc(self, u, lam, nt_cell=None):
        """solve_host through the allocation-free entry point (device staging
        in the library's fixed per-GPU array). Returns (status, iters)."""
        n = u.shape[0]
        assert u.flags.c_contiguous and lam.flags.c_contiguous
        status = np.zeros(n, dtype=np.int32)
        iters = np.zeros(n, dtype=np.int32)
        ntc = None if nt_cell is None else np.ascontiguousarray(nt_cell, dtype=np.int32)
        self._check(lib().uqk_closure_solve_host_noalloc(self.h, n, _p(ntc), _p(u), _p(lam), _p(status), _p(iters)),
                    "uqk_closure_solve_host_noalloc")
        return status, iters

    def eval_uq(self, lam, uq, nt_cell=None):
        self._check(lib().uqk_eval_uq(self.h, lam.shape[0], _p(nt_cell), _p(lam), _p(uq)), "uqk_eval_uq")

    def reconstruct(self, uq, u, nt_cell=None):
        self._check(lib().uqk_reconstruct(self.h, uq.shape[0], _p(nt_cell), _p(uq), _p(u)), "uqk_reconstruct")

    def compute_dt(self, prob, mesh, uq):
        dt = C.c_double()
        self._check(lib().uqk_compute_dt(self.h, C.byref(prob), C.byref(mesh.desc), _p(uq), C.byref(dt)),
                    "uqk_compute_dt")
        return dt.value

    def advance(self, prob, mesh, dt, u, uq, unew, nt_cell=None):
        self._check(lib().uqk_advance(self.h, C.byref(prob), C.byref(mesh.desc), _p(nt_cell), C.c_double(dt), _p(u),
                                      _p(uq), _p(unew)), "uqk_advance")

    def step(self, prob, mesh, u, lam, uq, u_scratch, status=None, iters=None, nt_cell=None):
        dt = C.c_double()
        res = C.c_double()
        nfail = self._check(lib().uqk_step(self.h, C.byref(prob), C.byref(mesh.desc), _p(nt_cell), _p(u), _p(lam),
                                           _p(uq), _p(u_scratch), _p(status), _p(iters), C.byref(dt), C.byref(res)),
                            "uqk_step")
        return nfail, dt.value, res.value

    def run_steps(self, prob, mesh, u, lam, uq, u_scratch, status=None, iters=None, nt_cell=None, t0=0.0,
                  t_end=1e300, max_steps=-1, use_graph=True):
        t, steps, res, newton = C.c_double(), C.c_int64(), C.c_double(), C.c_int64()
        nfail = self._check(lib().uqk_run_steps(self.h, C.byref(prob), C.byref(mesh.desc), _p(nt_cell), _p(u), _p(lam),
                                                _p(uq), _p(u_scratch), _p(status), _p(iters), t0, t_end, max_steps,
                                                int(use_graph), C.byref(t), C.byref(steps), C.byref(res),
                                                C.byref(newton)), "uqk_run_steps")
        return dict(nfail=nfail, t=t.value, steps=steps.value, residual=res.value, newton_steps=newton.value)

    def mean_var(self, lam, mv, nt_cell=None):
        self._check(lib().uqk_mean_var(self.h, lam.shape[0], _p(nt_cell), _p(lam), _p(mv)), "uqk_mean_var")

    def set_refinement(self, ntotal_per_level, refine_above=0.02, coarsen_below=0.004):
        """Refinement levels (momentsolver.cpp:104-117); step and run_steps then
        update nt_cell in place every step.  One level or None turns it off."""
        lv = np.ascontiguousarray([] if ntotal_per_level is None else ntotal_per_level, dtype=np.int32)
        self._check(lib().uqk_set_refinement(self.h, len(lv), _p(lv) if len(lv) else None, float(refine_above),
                                             float(coarsen_below)), "uqk_set_refinement")

    def set_refinement_cap(self, max_level):
        """Refinement retardation: no cell above level max_level."""
        self._check(lib().uqk_set_refinement_cap(self.h, int(max_level)), "uqk_set_refinement_cap")

    def refine(self, u, nt_cell):
        """One refinement pass: nt_cell (device int32, ncell) from the moments u."""
        self._check(lib().uqk_refine(self.h, u.shape[0], _p(u), _p(nt_cell)), "uqk_refine")

    def last_closure_ms(self):
        return lib().uqk_last_closure_ms(self.h)

    def last_closure_stats(self):
        a, b = C.c_int64(), C.c_int64()
        lib().uqk_last_closure_stats(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_factor_reuse(self, on=True):
        """Mixed closure: keep each cell's inverse Hessian and take refined
        steps on it until the contraction weakens (on by default for MIXED)."""
        self._check(lib().uqk_set_factor_reuse(self.h, int(bool(on))), "uqk_set_factor_reuse")

    def set_residual_check(self, on):
        """Residual check before the Newton kernel in the time loop (on by
        default): finishes the cells already converged, lists the others."""
        self._check(lib().uqk_set_residual_check(self.h, int(bool(on))), "uqk_set_residual_check")

    def invalidate_factors(self):
        """Forget the kept inverses (the next closure factors every cell)."""
        self._check(lib().uqk_invalidate_factors(self.h), "uqk_invalidate_factors")

    def last_closure_factorizations(self):
        """(Cholesky factorisations, steps on a kept inverse) of the last call or run."""
        a, b = C.c_int64(), C.c_int64()
        lib().uqk_last_closure_factorizations(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_kernel_timing(self, on=True):
        """Per-step HIP-event timing of the closure and advance kernels of
        steps enqueued without a graph; resets the sums."""
        self._check(lib().uqk_set_kernel_timing(self.h, int(bool(on))), "uqk_set_kernel_timing")

    def kernel_times(self):
        """(closure ms, advance ms, steps) summed since timing was turned on."""
        c, a, n = C.c_double(), C.c_double(), C.c_int64()
        lib().uqk_kernel_times(self.h, C.byref(c), C.byref(a), C.byref(n))
        return c.value, a.value, n.value


def problem(kind, flux=FLUX_REFERENCE, cfl=0.5, gamma=1.4, gravity=9.81, friction=0.0, friction_dim=-1):
    return ProblemDesc(kind, flux, cfl, gamma, gravity, friction, friction_dim)


class DeviceMesh:
    """Mesh arrays uploaded to device buffers (kept alive here)."""

    def __init__(self, mesh):
        self.nbr = DeviceArray.from_host(mesh.nbr, np.int32)
        self.unit_normal = DeviceArray.from_host(mesh.unit_normal, np.float64)
        self.normal = DeviceArray.from_host(mesh.normal, np.float64)
        self.area = DeviceArray.from_host(mesh.area, np.float64)
        self.dx_param = DeviceArray.from_host(mesh.dx_param, np.float64)
        self.bc = DeviceArray.from_host(mesh.bc, np.int32)
        self.bnormal = DeviceArray.from_host(mesh.bnormal, np.float64)
        ncell, nedge = mesh.nbr.shape
        self.desc = MeshDesc(ncell, nedge, mesh.unit_normal.shape[-1], _p(self.nbr), _p(self.unit_normal),
                             _p(self.normal), _p(self.area), _p(self.dx_param), _p(self.bc), _p(self.bnormal))


def synth_closure_euler1d(ncell, basis, seed=20261202, device=True):
    """The batched-closure workload (see include/uqk.h).  Returns (u, lam_star,
    lam0) as DeviceArrays (device=True) or numpy arrays, each (ncell, M, 3)."""
    M, Q = basis.M, basis.Q
    phi = np.ascontiguousarray(basis.phi)
    phiw = np.ascontiguousarray(basis.phiw)
    if device:
        out = [DeviceArray((ncell, M, 3)) for _ in range(3)]
        nf = lib().uqk_synth_euler1d_closure(ncell, seed, M, Q, _p(phi), _p(phiw), *[_p(a) for a in out])
    else:
        out = [np.zeros((ncell, M, 3)) for _ in range(3)]
        nf = lib().uqk_synth_euler1d_closure_host(ncell, seed, M, Q, _p(phi), _p(phiw), *[_p(a) for a in out])
    if nf != 0:
        raise RuntimeError("workload generation failed (%d)" % nf)
    return tuple(out)


def config_setup(path):
    """Host-only part of a run from an input file: (info, u, lambda, center)
    with info = dict(closure, S, M, Q, d, edges, dim, precision)."""
    err = C.create_string_buffer(512)
    info = np.zeros(8, np.int32)
    n = lib().uqk_config_setup(path.encode(), _p(info), None, None, None, err, 512)
    if n < 0:
        raise RuntimeError(err.value.decode())
    keys = ["closure", "S", "M", "Q", "d", "edges", "dim", "precision"]
    inf = dict(zip(keys, (int(v) for v in info)))
    u = np.zeros((n, inf["M"], inf["S"]))
    lam = np.zeros_like(u)
    center = np.zeros((n, inf["dim"]))
    n2 = lib().uqk_config_setup(path.encode(), None, _p(u), _p(lam), _p(center), err, 512)
    if n2 != n:
        raise RuntimeError(err.value.decode())
    return inf, u, lam, center


def run_config(path, mean_var_csv=None, output_prefix=None):
    """Runs an input file end to end on the GPU (uqk_run_config_output);
    returns the run report as a dict.  output_prefix: write <prefix>.vtk, the
    run log <prefix> and the restart files <prefix>_moments / _duals."""
    err = C.create_string_buffer(1024)
    rep = RunReport()
    rc = lib().uqk_run_config_output(path.encode(), mean_var_csv.encode() if mean_var_csv else None,
                                     output_prefix.encode() if output_prefix else None, C.byref(rep), err, 1024)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return {k: getattr(rep, k) for k, _ in RunReport._fields_}


class WarmStartNet:
    """Learned warm start of an Euler closure context (uqk_mlp_* entry points):
    four hidden layers of `hidden` GELU units, BF16 weights on MFMA, trained on
    the GPU with Adam on (u, lambda) pairs of device arrays."""

    def __init__(self, closure, hidden=256, seed=1):
        self.closure = closure
        h = C.c_void_p()
        if lib().uqk_mlp_create(closure.h, hidden, seed, C.byref(h)) != 0:
            raise RuntimeError("uqk_mlp_create: " + closure.error())
        self.h = h
        dims = np.zeros(6, np.int32)
        npar = C.c_int64()
        lib().uqk_mlp_shape(self.h, _p(dims), C.byref(npar))
        self.dims = [int(v) for v in dims]
        self.n_params = npar.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().uqk_mlp_destroy(self.h)
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, self.closure.error()))

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        nrm = np.zeros(2 * self.dims[0] + 2 * self.dims[-1], np.float32)
        self._check(lib().uqk_mlp_get_params(self.h, _p(p), _p(nrm)), "uqk_mlp_get_params")
        return p, nrm

    def set_params(self, params=None, norm=None):
        pp = None if params is None else np.ascontiguousarray(params, np.float32)
        nn = None if norm is None else np.ascontiguousarray(norm, np.float32)
        self._check(lib().uqk_mlp_set_params(self.h, _p(pp), _p(nn)), "uqk_mlp_set_params")

    def fit_normalization(self, u, lam):
        self._check(lib().uqk_mlp_fit_normalization(self.h, u.shape[0], _p(u), _p(lam)), "fit_normalization")

    def features(self, u, lam=None, standardise=False):
        n = u.shape[0]
        x = DeviceArray((n, self.dims[0]), np.float32)
        t = DeviceArray((n, self.dims[-1]), np.float32) if lam is not None else None
        self._check(lib().uqk_mlp_features(self.h, n, _p(u), _p(lam), int(standardise), _p(x), _p(t)),
                    "uqk_mlp_features")
        return x, t

    def forward(self, x):
        y = DeviceArray((x.shape[0], self.dims[-1]), np.float32)
        self._check(lib().uqk_mlp_forward(self.h, x.shape[0], _p(x), _p(y)), "uqk_mlp_forward")
        return y

    def loss_grad(self, x, t, n_real):
        g = DeviceArray(self.n_params, np.float32)
        loss = C.c_double()
        self._check(lib().uqk_mlp_loss_grad(self.h, x.shape[0], _p(x), _p(t), n_real, C.byref(loss), _p(g)),
                    "uqk_mlp_loss_grad")
        return loss.value, g

    def train(self, u, lam, batch=1024, steps=1000, lr=1e-3, seed=7):
        loss = C.c_double()
        self._check(lib().uqk_mlp_train(self.h, u.shape[0], _p(u), _p(lam), batch, steps, lr, seed,
                                        C.byref(loss)), "uqk_mlp_train")
        return loss.value

    def predict(self, u):
        out = DeviceArray(u.shape, np.float64)
        self._check(lib().uqk_mlp_predict(self.h, u.shape[0], _p(u), _p(out)), "uqk_mlp_predict")
        return out

    def attach(self):
        self._check(lib().uqk_set_warm_start(self.closure.h, self.h), "uqk_set_warm_start")

    def detach(self):
        lib().uqk_set_warm_start(self.closure.h, None)


def warm_start_select(closure, u, lam, lam_alt):
    n = lib().uqk_warm_start_select(closure.h, u.shape[0], _p(u), _p(lam), _p(lam_alt))
    if n < 0:
        raise RuntimeError("uqk_warm_start_select: " + closure.error())
    return n
