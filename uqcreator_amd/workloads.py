"""Synthetic workloads of the moment-solver hot path (host side, numpy).

closure_euler1d -- the batched-closure benchmark input: 1D Euler (gamma 1.4),
two uniform uncertain dimensions, total-degree-4 Legendre basis (M = 15),
7x7 Gauss quadrature (Q = 49).  Each cell's true duals lambda* are a
Maxwellian (rho in [0.5, 2], v in [-1, 1], p in [0.5, 2]) mapped through
EulerClosure::DS (eulerclosure.cpp:70-80) in the constant moment, plus
N(0, 0.05^2) in every higher moment.  A draw is kept only if the temperature
dual stays negative at every quadrature point, varies there by at most 3x,
and the density varies by at most 100x -- draws outside these bounds give
moment vectors at the edge of realizability whose residual cannot reach 1e-10
in FP64.  The moments are u = <U(lambda*^T phi) phi> in FP64, and the Newton
start is lambda_0 = (DS(u_0), 0, ...) as in MomentSolver::SetDuals
(momentsolver.cpp:473-484).  Random numbers come from a torch generator
seeded with 20261202 on the device the inputs are built on (on MI355X that is
the Philox-based device generator), so 2^20 cells take a second to build.
"""
import numpy as np

from . import native

SEED = 20261202
GAMMA = 1.4


def _euler1d_U(L, g=GAMMA):
    import torch

    v0, v1, v2 = L[..., 0], L[..., 1], L[..., 2]
    r = torch.exp((2 * v0 * v2 - 2 * v2 * torch.log(-v2) - 2 * v2 * g - v1 * v1) / (2 * v2 * (g - 1)))
    return torch.stack([r, -(v1 / v2) * r, ((v1 * v1 - 2 * v2) / (2 * v2 * v2)) * r], -1)


def euler1d_DS(u, g=GAMMA):
    import torch

    rho, ru, rE = u[..., 0], u[..., 1], u[..., 2]
    gm = -g
    ru2 = ru * ru
    d0 = (ru2 + gm * (2 * rho * rE - ru2)) / (-2 * rho * rE + ru2) - torch.log(rho ** gm * (rE - ru2 / (2 * rho)))
    return torch.stack([d0, -(2 * rho * ru) / (-2 * rho * rE + ru2), -(rho / (rE - ru2 / (2 * rho)))], -1)


def closure_basis():
    return native.Basis(d=2, degrees=[4], q1=7)


def closure_euler1d(ncell, basis=None, seed=SEED, temp_ratio=1.0 / 3.0, rho_ratio=100.0, device="cpu",
                    batch=1 << 20):
    """Returns (u, lam_star, lam0) as float64 torch tensors (ncell, M, S=3) on
    `device`.  Candidates are drawn in batches from one torch generator seeded
    with `seed` on that device and accepted in draw order, so the result depends
    only on (seed, ncell, device type)."""
    import torch

    b = basis or closure_basis()
    dt = torch.float64
    phi = torch.as_tensor(b.phi, dtype=dt, device=device)
    phiw = torch.as_tensor(b.phiw, dtype=dt, device=device)
    M = b.M
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    lam = torch.zeros((ncell, M, 3), dtype=dt, device=device)
    filled = 0
    while filled < ncell:
        n = batch
        r3 = torch.rand((n, 3), generator=gen, dtype=dt, device=device)
        rho = 0.5 + 1.5 * r3[:, 0]
        v = -1.0 + 2.0 * r3[:, 1]
        p = 0.5 + 1.5 * r3[:, 2]
        pert = 0.05 * torch.randn((n, M - 1, 3), generator=gen, dtype=dt, device=device)
        cons = torch.stack([rho, rho * v, p / (GAMMA - 1) + 0.5 * rho * v * v], -1)
        l = torch.cat([euler1d_DS(cons)[:, None, :], pert], 1)
        L = torch.einsum("qm,nms->nqs", phi, l)
        L2 = L[..., 2]
        ok = (L2.amax(1) < 0) & (L2.amax(1) / L2.amin(1) >= temp_ratio)
        r = _euler1d_U(torch.where(ok[:, None, None], L, -torch.ones_like(L)))[..., 0]
        ok &= r.amax(1) / r.amin(1) <= rho_ratio
        l = l[ok][: ncell - filled]
        lam[filled:filled + l.shape[0]] = l
        filled += l.shape[0]
    u = torch.empty_like(lam)
    lam0 = torch.zeros_like(lam)
    for s in range(0, ncell, batch):
        L = torch.einsum("qm,nms->nqs", phi, lam[s:s + batch])
        u[s:s + batch] = torch.einsum("qm,nqs->nms", phiw, _euler1d_U(L))
        lam0[s:s + batch, 0] = euler1d_DS(u[s:s + batch, 0])
    return u, lam, lam0


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
