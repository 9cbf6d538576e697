"""Synthetic workloads of the moment-solver hot path (host side).

closure_euler1d -- the batched-closure benchmark input: 1D Euler (gamma 1.4),
two uniform uncertain dimensions, total-degree-4 Legendre basis (M = 15),
7x7 Gauss quadrature (Q = 49).  Each cell's true duals lambda* are a
Maxwellian (rho in [0.5, 2], v in [-1, 1], p in [0.5, 2]) mapped through
EulerClosure::DS (eulerclosure.cpp:70-80) in the constant moment, plus
N(0, 0.05^2) in every higher moment.  A draw is kept only if the temperature
dual stays negative at every quadrature point and varies there by at most 3x,
and the density varies by at most 100x -- draws outside these bounds give
moment vectors at the edge of realizability whose residual cannot reach 1e-10
in FP64.  The moments are u = <U(lambda*^T phi) phi> in FP64 and the Newton
start is lambda_0 = (DS(u_0), 0, ...) as in MomentSolver::SetDuals
(momentsolver.cpp:473-484).  Random numbers come from a counter-based
Philox4x32-10 stream keyed by the seed with counter (cell, attempt, draw), so
each cell is reproducible on its own, on the host and on the device
(uqk_synth_euler1d_closure in the native library).
"""
import numpy as np

from . import native

SEED = 20261202
GAMMA = 1.4


def closure_basis():
    return native.Basis(d=2, degrees=[4], q1=7)


def closure_euler1d(ncell, basis=None, seed=SEED, device=False):
    """Returns (u, lam_star, lam0), each (ncell, M, 3) float64: numpy arrays
    (device=False) or DeviceArrays (device=True)."""
    b = basis or closure_basis()
    return native.synth_closure_euler1d(ncell, b, seed=seed, device=device)


def project_ic(ic, centers, basis):
    """SetupIC (momentsolver.cpp:535-563): u_j = IC(x_j, xi_k) phiw.
    ic(x (ncell, dim), xi (Q, d)) -> (ncell, Q, S)."""
    vals = ic(centers, basis.xi)
    return np.ascontiguousarray(np.einsum("qm,nqs->nms", basis.phiw, vals))


def burgers_step_ic(x0=0.5, sigma=0.2, uL=12.0, uR=3.0):
    """1D Burgers shock: uL left of x0 + sigma xi, uR right of it."""

    def ic(x, xi):
        shock = x0 + sigma * xi[:, 0]
        return np.where(x[:, 0:1] < shock[None, :], uL, uR)[..., None]

    return ic


def burgers_c1(n_cells=2000):
    """Config: 1D Burgers, bounded barrier [2.9, 12.1], N = 5 (6 moments),
    12 Gauss points, LF flux, CFL 0.5, t_end 0.11, on [0, 3]."""
    from . import mesh as _mesh

    b = native.Basis(1, [5], 12)
    m = _mesh.mesh1d(n_cells, 0.0, 3.0)
    u = project_ic(burgers_step_ic(), m.center, b)
    lam = np.zeros_like(u)  # BoundedBarrier has no DS: lambda starts at 0 (closure.cpp:289)
    return b, m, u, lam


def euler2d_freestream_ic(gamma=1.4, R=287.87, T=273.15, p=101325.0, mach=0.8, dmach=0.02, aoa=1.25, daoa=0.5):
    """Uncertain free stream around the profile (the reference's NACA branch,
    euler2d.cpp:182-206): Mach = mach + dmach xi_2, angle of attack
    aoa + daoa xi_1 degrees, SI units, the same state in every cell."""

    def ic(x, xi):
        a = np.sqrt(gamma * R * T)
        ma = mach + (dmach * xi[:, 1] if xi.shape[1] > 1 else 0.0)
        ang = (aoa + daoa * xi[:, 0]) * (2.0 * np.pi) / 360.0
        rho = p / (R * T)
        uf, vf = ma * a * np.cos(ang), ma * a * np.sin(ang)
        st = np.stack([np.full_like(uf, rho), rho * uf, rho * vf,
                       0.5 * rho * (uf ** 2 + vf ** 2) + p / (gamma - 1)], -1)
        return np.broadcast_to(st[None], (x.shape[0],) + st.shape).copy()

    return ic


def euler2d_c3(ni=512, nj=128):
    """Config: 2D Euler past NACA0012, O-grid ni x nj, total degree 3 in two
    uncertain dims (M = 10), 5x5 Gauss (Q = 25)."""
    from . import mesh as _mesh

    b = native.Basis(2, [3], 5)
    m = _mesh.naca_ogrid(ni, nj)
    u = project_ic(euler2d_freestream_ic(), m.center, b)
    return b, m, u


def sw2d_dambreak_ic(h_right=1.0, ratio=2.0, dratio=0.5, dpos=0.1):
    """Dam break: depth (ratio + dratio xi_1) h_right left of x = dpos xi_2,
    h_right to the right, water at rest."""

    def ic(x, xi):
        hl = (ratio + dratio * xi[:, 0]) * h_right
        dam = dpos * xi[:, 1]
        h = np.where(x[:, 0:1] < dam[None, :], hl[None, :], h_right)
        return np.stack([h, np.zeros_like(h), np.zeros_like(h)], -1)

    return ic


def sw2d_c5(n=1024, degree=4, q1=5):
    """Config: 2D shallow-water dam break on [-1,1]^2 (n x n cells, Dirichlet
    frame), three uncertain dims, total degree 4 (M = 35), 5^3 Gauss points."""
    from . import mesh as _mesh

    b = native.Basis(3, [degree], q1)
    m = _mesh.quad_grid(n, n, -1.0, 1.0, -1.0, 1.0, native.BC_DIRICHLET)
    u = project_ic(sw2d_dambreak_ic(), m.center, b)
    return b, m, u


def cell_orders(n_cells, seed=SEED, n_max=6):
    """Per-cell moment order N in {1..n_max}, uniform, from numpy's Philox
    generator keyed by the seed: the uneven per-cell Newton sizes of an
    adaptive run, reproducible from the seed."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    return 1 + rng.integers(0, n_max, size=n_cells)


def euler2d_c4(ni=512, nj=128, seed=SEED):
    """Config: the 2D Euler NACA0012 setup with per-cell orders N in 1..6
    (graded total-degree levels, n_total = 3, 6, 10, 15, 21, 28), 7x7 Gauss.
    Returns (basis, mesh, u, lam0, nt_cell)."""
    from . import mesh as _mesh

    b = native.Basis(2, [1, 2, 3, 4, 5, 6], 7)
    m = _mesh.naca_ogrid(ni, nj)
    u = project_ic(euler2d_freestream_ic(), m.center, b)
    orders = cell_orders(m.n_cells, seed)
    nt = np.asarray(b.ntotal_per_level, dtype=np.int32)[orders - 1]
    lam0 = np.zeros_like(u)
    lam0[:, 0] = euler2d_ds(u[:, 0])
    return b, m, u, lam0, nt


def euler2d_ds(u0, gamma=1.4):
    """EulerClosure2D::DS (eulerclosure2d.cpp:90-102) of the mean states
    u0 (n, 4): the plain Newton start lambda_0 = (DS(u_0), 0, ...)."""
    g = -gamma
    rho, ru, rv, rE = u0[:, 0], u0[:, 1], u0[:, 2], u0[:, 3]
    q2 = ru * ru + rv * rv
    out = np.empty_like(u0)
    out[:, 0] = (q2 + g * (2 * rho * rE - q2)) / (-2 * rho * rE + q2) - np.log(rho ** g * (rE - q2 / (2 * rho)))
    out[:, 1] = -((2 * rho * ru) / (-2 * rho * rE + q2))
    out[:, 2] = -((2 * rho * rv) / (-2 * rho * rE + q2))
    out[:, 3] = -(rho / (rE - q2 / (2 * rho)))
    return out
