"""Host side of the product: the C ABI library loads and exports every symbol
of include/uqk.h; the host numerics (quadrature, index set, Closure tables)
match the library fixtures and the oracle.  No GPU work."""
import os
import re

import numpy as np
import pytest

from uqcreator_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "uqk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(uqk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = native.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(native.SIGNATURES), set(syms) ^ set(native.SIGNATURES)


def test_gauss_legendre_golub_welsch(golden):
    for n in [1, 2, 3, 5, 7, 12, 15, 20, 25, 49]:
        x, w = native.gauss_legendre(n)
        np.testing.assert_allclose(x, golden["gl_x_%d" % n], atol=4e-15)
        np.testing.assert_allclose(w, golden["gl_w_%d" % n], rtol=1e-12 if n < 20 else 1e-11, atol=1e-15)


def test_host_tables_match_fixture_and_oracle(golden, oracle):
    b = native.Basis(2, [4], 7)
    np.testing.assert_array_equal(b.idx, golden["tab_idx"])
    np.testing.assert_allclose(b.phi, golden["tab_phi"], atol=1e-13)
    np.testing.assert_allclose(b.phiw, golden["tab_phiw"], atol=1e-14)
    for d, q, degs in [(1, 12, [5]), (2, 5, [3]), (3, 5, [4]), (2, 7, [1, 2, 3, 4, 5, 6])]:
        b = native.Basis(d, degs, q)
        idx, ntl = oracle.index_set(d, degs)
        np.testing.assert_array_equal(b.idx, idx)
        np.testing.assert_array_equal(b.ntotal_per_level, ntl)
        _, _, phi, phiw = oracle.tables(d, q, idx)
        np.testing.assert_allclose(b.phi, phi, atol=1e-12)
        np.testing.assert_allclose(b.phiw, phiw, atol=1e-13)


def test_context_rejects_bad_descriptors():
    b = native.Basis(1, [5], 12)
    with pytest.raises(RuntimeError):
        native.GpuClosure(native.EULER2D, b, precision=7)


def test_workload_generator_is_consistent(oracle):
    from uqcreator_amd import workloads

    b = workloads.closure_basis()
    u, lam, lam0 = workloads.closure_euler1d(32, basis=b)
    c = oracle.closure(1, du_exact=1)
    for j in range(32):
        # u are the moments of the closure at lambda*
        assert np.linalg.norm(oracle.gradient(c, b.phi, b.phiw, lam[j], u[j])) < 1e-13
        np.testing.assert_allclose(lam0[j, 0], oracle.DS(c, u[j, 0]), rtol=1e-12)
    u2, lam2, _ = workloads.closure_euler1d(32, basis=b)
    np.testing.assert_array_equal(lam, lam2)  # seeded
