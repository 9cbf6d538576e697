"""Writes tests/golden/fixtures.npz: known-answer vectors computed with numpy /
scipy library routines, independent of both the oracle and the product.

    python tests/golden/make_golden.py

Contents
  gl_x_<n>, gl_w_<n>   numpy.polynomial.legendre.leggauss(n), n in GL_NS
  leg_x, leg_<n>       scipy.special.eval_legendre(n, leg_x), n = 0..12
  tab_*                tensor Gauss-Legendre tables for d = 2, q = 7, total
                       degree 4, built from leggauss + legval with the
                       multi-index order of settings.cpp:242-266
  chol_H, chol_b,      an SPD matrix, its upper Cholesky factor (scipy
  chol_U, chol_x       cho_factor) and cho_solve(H, b)
"""
import os

import numpy as np
import scipy.linalg as sl
import scipy.special as sp

GL_NS = list(range(1, 16)) + [20, 25, 49]


def total_degree_indices(d, N):
    out = []
    for i in range((N + 1) ** d):
        mi = [(i // (N + 1) ** l) % (N + 1) for l in range(d)]
        if sum(mi) <= N:
            out.append(mi)
    return np.array(out, dtype=np.int32)


def main():
    out = {}
    for n in GL_NS:
        x, w = np.polynomial.legendre.leggauss(n)
        out["gl_x_%d" % n] = x
        out["gl_w_%d" % n] = w
    lx = np.linspace(-1, 1, 41)
    out["leg_x"] = lx
    for n in range(13):
        out["leg_%d" % n] = sp.eval_legendre(n, lx)
    d, q, N = 2, 7, 4
    idx = total_degree_indices(d, N)
    x1, w1 = np.polynomial.legendre.leggauss(q)
    Q = q ** d
    xi = np.zeros((Q, d))
    wq = np.ones(Q)
    for k in range(Q):
        for l in range(d):
            a = (k // q ** l) % q
            xi[k, l] = x1[a]
            wq[k] *= w1[a]
    M = len(idx)
    phi = np.ones((Q, M))
    phiw = np.ones((Q, M))
    for i, mi in enumerate(idx):
        for l in range(d):
            c = np.zeros(mi[l] + 1)
            c[-1] = 1.0
            v = np.polynomial.legendre.legval(xi[:, l], c) * np.sqrt(2 * mi[l] + 1)
            phi[:, i] *= v
            phiw[:, i] *= v * 0.5
        phiw[:, i] *= wq
    out.update(tab_idx=idx, tab_xi=xi, tab_wq=wq, tab_phi=phi, tab_phiw=phiw)
    rng = np.random.default_rng(7)
    A = rng.normal(size=(12, 12))
    H = A @ A.T + 12 * np.eye(12)
    b = rng.normal(size=12)
    c, low = sl.cho_factor(H, lower=False)
    out.update(chol_H=H, chol_b=b, chol_U=np.triu(c), chol_x=sl.cho_solve((c, low), b))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.npz")
    np.savez(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
