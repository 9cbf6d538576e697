"""Refinement levels of the adaptive moment solver (momentsolver.cpp:104-117,
ComputeRefIndicator :271-289) in the oracle, on the CPU.

The indicator is checked against a characterisation that does not use the
nTotal prefix arithmetic: the moments a level adds are the ones whose total
degree lies above the previous level's degree (the index set is graded by
level), and with one uncertain dimension the indicator takes the moments of
the two highest degrees of the level."""
import numpy as np
import pytest

from uqcreator_amd import mesh as meshes
from uqcreator_amd import native, workloads

from test_gpu_sweeps import _oracle_mesh


@pytest.mark.parametrize("d,degrees", [(1, [2, 3, 4, 5]), (2, [1, 2, 3]), (2, [0, 2, 4]), (3, [1, 3])])
def test_indicator_by_degree(oracle, d, degrees):
    idx, ntl = native.index_set(d, degrees)
    tdeg = idx.sum(1)
    rng = np.random.default_rng(7)
    S = 3
    u = rng.normal(size=(len(idx), S))
    for lev, deg in enumerate(degrees):
        got = oracle.ref_indicator(S, d, ntl, lev, u)
        if d == 1:
            sel = (tdeg == deg) | (tdeg == deg - 1)
        else:
            lo = degrees[lev - 1] if lev > 0 else 0
            sel = (tdeg > lo) & (tdeg <= deg)
        want = np.abs(u[sel, 0]).sum()
        assert got == pytest.approx(want, rel=1e-15, abs=0.0), (lev, got, want)


def test_refine_rule(oracle):
    """up above 0.02 unless finest, down below 0.004 unless coarsest, else stay;
    a cell whose nTotal is no level's is left alone"""
    d, degrees = 2, [1, 2, 3]
    idx, ntl = native.index_set(d, degrees)  # nTotal 3, 6, 10
    M, S = len(idx), 4
    cases = [  # (level, indicator, expected level)
        (0, 0.03, 1), (1, 0.03, 2), (2, 0.03, 2),
        (0, 0.001, 0), (1, 0.001, 0), (2, 0.001, 1),
        (0, 0.01, 0), (1, 0.01, 1), (2, 0.01, 2),
        (1, 0.02, 1), (1, 0.004, 1),  # the bounds themselves stay
    ]
    u = np.zeros((len(cases) + 1, M, S))
    nt = np.zeros(len(cases) + 1, np.int32)
    for j, (lev, ind, _) in enumerate(cases):
        lo = ntl[lev - 1] if lev > 0 else 1
        u[j, lo, 0] = -ind  # the indicator takes |u|
        u[j, :, 1:] = 5.0  # other states do not count
        nt[j] = ntl[lev]
    nt[-1] = 4  # not a level
    u[-1, 1:, 0] = 1.0
    out = oracle.refine(u, d, ntl, nt)
    for j, (lev, ind, want) in enumerate(cases):
        assert out[j] == ntl[want], (lev, ind, out[j])
    assert out[-1] == 4


def test_refine_cap(oracle):
    """refinement retardation: no cell ends above the cap; one above it comes
    down to it whatever its indicator"""
    d, degrees = 2, [1, 2, 3]
    idx, ntl = native.index_set(d, degrees)
    M, S = len(idx), 4
    cases = [(0, 0.03, 1), (1, 0.03, 1), (2, 0.03, 1), (2, 0.01, 1), (2, 0.001, 1), (1, 0.001, 0), (0, 0.01, 0)]
    u = np.zeros((len(cases), M, S))
    nt = np.zeros(len(cases), np.int32)
    for j, (lev, ind, _) in enumerate(cases):
        u[j, ntl[lev - 1] if lev > 0 else 1, 0] = ind
        nt[j] = ntl[lev]
    out = oracle.refine(u, d, ntl, nt, max_level=1)
    assert out.tolist() == [ntl[w] for _, _, w in cases]


def _burgers_levels(n_cells=200):
    b = native.Basis(1, [2, 3, 4, 5], 12)
    m = meshes.mesh1d(n_cells, 0.0, 3.0)
    u = workloads.project_ic(workloads.burgers_step_ic(), m.center, b)
    return b, m, u, np.zeros_like(u)


def test_run_levels_without_moves_is_the_plain_loop(oracle):
    """thresholds that never fire give the fixed-order loop bit for bit"""
    b, m, u0, lam0 = _burgers_levels()
    c = oracle.closure(oracle.BB, u_minus=2.9, u_plus=12.1, eps=1e-10)
    p = oracle.problem(oracle.P_BURGERS, flux=oracle.FLUX_LF, cfl=0.5)
    om = _oracle_mesh(oracle, m)
    ua, la = u0.copy(), lam0.copy()
    nt = np.full(m.n_cells, b.M, np.int32)
    ra = oracle.run_levels(c, p, om, b.phi, b.phiw, b.xi, ua, la, b.ntotal_per_level, nt, 1e3, 5, above=np.inf,
                           below=-np.inf, nthreads=1)
    ub, lb = u0.copy(), lam0.copy()
    rb = oracle.run(c, p, om, b.phi, b.phiw, b.xi, ub, lb, 1e3, 5, nthreads=1)
    assert ra == rb
    assert (nt == b.M).all()
    assert np.array_equal(ua, ub) and np.array_equal(la, lb)


def test_run_levels_moves_cells_and_truncates(oracle):
    """the Burgers step: away from the shock the higher moments vanish and the
    cells coarsen to level 0; at the shock they stay fine; the moments above
    a cell's nTotal are zero after every step"""
    b, m, u, lam = _burgers_levels()
    c = oracle.closure(oracle.BB, u_minus=2.9, u_plus=12.1, eps=1e-10)
    p = oracle.problem(oracle.P_BURGERS, flux=oracle.FLUX_LF, cfl=0.5)
    nt = np.full(m.n_cells, b.M, np.int32)
    r = oracle.run_levels(c, p, _oracle_mesh(oracle, m), b.phi, b.phiw, b.xi, u, lam, b.ntotal_per_level, nt, 1e3, 6,
                          nthreads=2)
    assert r["rc"] == 0 and r["steps"] == 6
    ntl = list(b.ntotal_per_level)
    assert set(nt.tolist()) <= set(ntl)
    assert (nt == ntl[0]).sum() > m.n_cells // 2
    x = m.center[:, 0]
    near = np.abs(x - 0.5) < 0.05
    assert (nt[near] == ntl[-1]).any()
    for j in range(m.n_cells):
        assert not u[j, nt[j]:].any(), j
