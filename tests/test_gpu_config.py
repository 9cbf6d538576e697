"""Whole runs from input files through uqk_run_config (the C++ host layer:
Settings, Mesh, SetupIC, SetDuals, the device time loop, WriteMeanAndVar)
against the oracle's time loop on the same initial state."""
import os

import numpy as np
import pytest

from uqcreator_amd import mesh as meshes
from uqcreator_amd import native

from test_gpu_sweeps import _oracle_closure, _oracle_mesh, _rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUT = os.path.join(ROOT, "uqcreator_amd", "input")

pytestmark = pytest.mark.gpu


def _run(oracle, path, tmp_path, basis, m, kind, prob, clo, t_end, max_steps):
    csv = str(tmp_path / "mv.csv")
    print("case: input", path, flush=True)
    rep = native.run_config(path, csv)
    inf, u0, lam0, center = native.config_setup(path)
    assert (inf["M"], inf["Q"]) == (basis.M, basis.Q)
    np.testing.assert_allclose(center, m.center, atol=1e-12)
    mv = np.loadtxt(csv, delimiter=",", ndmin=2)
    c = _oracle_closure(oracle, kind, **clo)
    res = oracle.run(c, oracle.problem(**prob), _oracle_mesh(oracle, m), basis.phi, basis.phiw, basis.xi, u0, lam0,
                     t_end, max_steps)
    assert res["rc"] == 0
    assert rep["steps"] == res["steps"], (rep, res)
    mv_o = oracle.mean_var(c, basis.phi, basis.phiw, lam0)
    return rep, _rel_l2(mv, mv_o, m.area)


def test_burgers_input_end_to_end(oracle, tmp_path):
    rep, err = _run(oracle, os.path.join(INPUT, "burgers1d_step.toml"), tmp_path, native.Basis(1, [5], 12),
                    meshes.mesh1d(2000, 0.0, 3.0), native.BOUNDED_BARRIER,
                    dict(kind=native.BURGERS, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.5),
                    dict(u_minus=2.9, u_plus=12.1, eps=1e-10), 0.11, -1)
    print("report", rep, "rel L2 E/Var", err, flush=True)
    assert rep["steps"] > 1500 and rep["t"] >= 0.11 - 1e-12
    assert err.max() < 1e-8


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("mixed", 1e-6)])
def test_euler2d_input_short_run(oracle, tmp_path, precision, tol):
    text = open(os.path.join(INPUT, "euler2d_naca.toml")).read()
    for a, b in [("ni = 512", "ni = 64"), ("nj = 128", "nj = 16"), ("maxSteps = 2000", "maxSteps = 10"),
                 ('precision = "fp64"', 'precision = "%s"' % precision)]:
        assert a in text
        text = text.replace(a, b)
    path = tmp_path / "naca.toml"
    path.write_text(text)
    rep, err = _run(oracle, str(path), tmp_path, native.Basis(2, [3], 5), meshes.naca_ogrid(64, 16), native.EULER2D,
                    dict(kind=native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.4), dict(eps=1e-6),
                    1e3, 10)
    print(precision, "report", rep, "rel L2 E/Var", err, flush=True)
    assert rep["steps"] == 10
    assert err.max() < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("mixed", 1e-6)])
def test_shallow_water_input_short_run(oracle, tmp_path, precision, tol):
    """The dam-break input (barrier closure, n = 105, 5^3 points, bed friction
    on the third uncertain dimension) on a 16 x 16 grid, 4 steps, against the
    oracle's time loop, with the closure in FP64 and in the input's mixed
    precision (Hessian tiles on FP32 MFMA in the workgroup kernel)."""
    text = open(os.path.join(INPUT, "shallowwater2d_dambreak.toml")).read()
    for a, b in [("nx = 1024", "nx = 16"), ("ny = 1024", "ny = 16"), ("maxSteps = 500", "maxSteps = 4"),
                 ('precision = "mixed"', 'precision = "%s"' % precision)]:
        assert a in text
        text = text.replace(a, b)
    path = tmp_path / "sw.toml"
    path.write_text(text)
    rep, err = _run(oracle, str(path), tmp_path, native.Basis(3, [4], 5),
                    meshes.quad_grid(16, 16, -1.0, 1.0, -1.0, 1.0, native.BC_DIRICHLET),
                    native.SHALLOW_WATER2D_BARRIER,
                    dict(kind=native.SHALLOW_WATER2D_PROBLEM, flux=native.FLUX_REFERENCE, cfl=0.45, friction=0.01,
                         friction_dim=2),
                    dict(barrier=0.01, eps=1e-8), 1e3, 4)
    print(precision, "report", rep, "rel L2 E/Var", err, flush=True)
    assert rep["steps"] == 4
    assert err.max() < tol


def test_euler2d_su2_triangles_input(oracle, tmp_path):
    """Euler 2D on an SU2 triangle mesh (the 64 x 16 O-grid nodes split into
    2048 triangles, profile 'noslip', far field 'dirichlet') through an input
    file with format = "SU2", 10 steps, against the oracle's loop on the mesh
    the same reader returns."""
    meshes.naca_ogrid_su2(str(tmp_path / "naca_tri.su2"), 64, 16)
    text = open(os.path.join(INPUT, "euler2d_naca.toml")).read()
    for a, b in [('type = "nacaOGrid"', 'format = "SU2"\nSU2MeshFile = "naca_tri.su2"\n'
                  'SU2BC = [["farfield", "dirichlet"], ["airfoil", "noslip"]]'), ("maxSteps = 2000", "maxSteps = 10")]:
        assert a in text
        text = text.replace(a, b)
    path = tmp_path / "naca_su2.toml"
    path.write_text(text)
    m = meshes.su2(str(tmp_path / "naca_tri.su2"), {"farfield": "dirichlet", "airfoil": "noslip"})
    rep, err = _run(oracle, str(path), tmp_path, native.Basis(2, [3], 5), m, native.EULER2D,
                    dict(kind=native.EULER2D_PROBLEM, flux=native.FLUX_LAX_FRIEDRICHS, cfl=0.4), dict(eps=1e-6),
                    1e3, 10)
    print("report", rep, "rel L2 E/Var", err, flush=True)
    assert rep["steps"] == 10 and rep["n_cells"] == 2048
    assert err.max() < 1e-8


def _read_vtk(path):
    """A legacy ASCII VTK unstructured grid: points, cells, types and the
    FIELD arrays of CELL_DATA and POINT_DATA."""
    tok = open(path).read().split()
    i = tok.index("POINTS")
    P = int(tok[i + 1])
    pts = np.array(tok[i + 3:i + 3 + 3 * P], float).reshape(P, 3)
    i = tok.index("CELLS")
    N = int(tok[i + 1])
    j, cells = i + 3, []
    for _ in range(N):
        k = int(tok[j])
        cells.append([int(v) for v in tok[j + 1:j + 1 + k]])
        j += k + 1
    i = tok.index("CELL_TYPES")
    types = [int(v) for v in tok[i + 2:i + 2 + N]]
    data = {}
    for sect in ("CELL_DATA", "POINT_DATA"):
        i = tok.index(sect)
        nf = int(tok[i + 4])
        j = i + 5
        for _ in range(nf):
            name, comp, n = tok[j], int(tok[j + 1]), int(tok[j + 2])
            data[(sect, name)] = np.array(tok[j + 4:j + 4 + comp * n], float).reshape(n, comp)
            j += 4 + comp * n
    return pts, cells, types, data


def _read_field(path, n, M, S):
    rows = [l.rstrip("\n").split(",") for l in open(path)]
    assert len(rows) == n and all(r[-1] == "" for r in rows)
    x = np.array([r[:-1] for r in rows], float).reshape(n, S, M)
    return np.ascontiguousarray(x.transpose(0, 2, 1))


def test_output_files_and_restart(oracle, tmp_path):
    """Euler 2D on a 32 x 8 O-grid: a run to tEnd 2e-5 writes the fields
    (VTK cell data equal to the mean/variance, point data their averages over
    each node's cells), the run log and the restart files; an input with
    restartFile = that run and importDualState continues from tEnd for 5
    steps and matches the oracle started from the same files."""
    text = open(os.path.join(INPUT, "euler2d_naca.toml")).read()
    for a, b in [("ni = 512", "ni = 32"), ("nj = 128", "nj = 8"), ("tEnd = 1e3", "tEnd = 2e-5")]:
        assert a in text
        text = text.replace(a, b)
    p1 = tmp_path / "run1.toml"
    p1.write_text(text)
    pre = str(tmp_path / "out" / "run1")
    os.makedirs(tmp_path / "out")
    csv = str(tmp_path / "mv1.csv")
    print("case: output files", flush=True)
    rep1 = native.run_config(str(p1), csv, pre)
    assert rep1["t"] >= 2e-5 and rep1["steps"] >= 4
    n, S, M = 256, 4, 10
    mv = np.loadtxt(csv, delimiter=",", ndmin=2)
    pts, cells, types, data = _read_vtk(pre + ".vtk")
    m = meshes.naca_ogrid(32, 8)
    assert len(cells) == n and set(types) == {9} and pts.shape == (32 * 9, 3)
    np.testing.assert_allclose(pts[np.array(cells)].mean(1)[:, :2], m.center, atol=1e-12)
    rho, mom = data[("CELL_DATA", "E(%CF%81)")], data[("CELL_DATA", "Var(%CF%81U)")]
    np.testing.assert_array_equal(rho[:, 0], mv[:, 0])
    np.testing.assert_array_equal(mom[:, :2], mv[:, 5:7])
    assert not mom[:, 2].any()
    users = np.zeros(len(pts))
    acc = np.zeros(len(pts))
    for c, nodes in enumerate(cells):
        users[nodes] += 1
        acc[nodes] += mv[c, 3]
    np.testing.assert_allclose(data[("POINT_DATA", "E(%CF%81E)")][:, 0], acc / users, rtol=1e-14)
    log = open(pre).read()
    assert "Config file:\t" + str(p1) in log and "ni = 32" in log
    u1 = _read_field(pre + "_moments", n, M, S)
    l1 = _read_field(pre + "_duals", n, M, S)

    # continue: restartFile + importDualState, tEnd 1e3, 5 more steps
    text2 = text.replace("tEnd = 2e-5", "tEnd = 1e3").replace("maxSteps = 2000", "maxSteps = 5").replace(
        '[general]', '[general]\nrestartFile = "out/run1"\nimportDualState = true')
    p2 = tmp_path / "run2.toml"
    p2.write_text(text2)
    inf, u0, lam0, _ = native.config_setup(str(p2))
    np.testing.assert_array_equal(u0, u1)
    np.testing.assert_array_equal(lam0, l1)
    csv2 = str(tmp_path / "mv2.csv")
    rep2 = native.run_config(str(p2), csv2)
    assert rep2["steps"] == 5 and rep2["t"] > 2e-5
    b = native.Basis(2, [3], 5)
    c = _oracle_closure(oracle, native.EULER2D, eps=1e-6)
    oprob = oracle.problem(oracle.P_EULER2D, flux=oracle.FLUX_LF, cfl=0.4)
    uo, lo = u1.copy(), l1.copy()
    res = oracle.run(c, oprob, _oracle_mesh(oracle, m), b.phi, b.phiw, b.xi, uo, lo, 1e3, 5)
    assert res["rc"] == 0 and res["steps"] == 5
    assert rep2["t"] == pytest.approx(2e-5 + res["t"], rel=1e-12)
    err = _rel_l2(np.loadtxt(csv2, delimiter=",", ndmin=2), oracle.mean_var(c, b.phi, b.phiw, lo), m.area)
    print("restart rel L2", err, flush=True)
    assert err.max() < 1e-8


def test_restart_raises_truncation_order(oracle, tmp_path):
    """A run at total degree 2 on 3 x 3 points restarted at total degree 3 on
    5 x 5 points (momentsolver.cpp:418-523): the start state equals the
    oracle's restatement of the reference's sequence (Dirichlet cells take
    the initial condition's first moments, DS start converged by the previous
    closure, Lambda of the previous basis at the new points, U there,
    projected on the new basis, duals padded with zeros and converged by the
    new closure), and 5 steps from it match the oracle."""
    text = open(os.path.join(INPUT, "euler2d_naca.toml")).read()
    for a, b in [("ni = 512", "ni = 32"), ("nj = 128", "nj = 8"), ("tEnd = 1e3", "tEnd = 2e-5"),
                 ('moments = [["totalDegree"], [3]]', 'moments = [["totalDegree"], [2]]'),
                 ('quadPoints = [["tensorizedGrid"], [5]]', 'quadPoints = [["tensorizedGrid"], [3]]')]:
        assert a in text
        text = text.replace(a, b)
    p1 = tmp_path / "run1.toml"
    p1.write_text(text)
    pre = str(tmp_path / "out" / "run1")
    os.makedirs(tmp_path / "out")
    print("case: restart raising the order, first run (degree 2, 3x3 points)", flush=True)
    rep1 = native.run_config(str(p1), None, pre)
    assert rep1["t"] >= 2e-5
    n, S, M1, M2 = 256, 4, 6, 10
    u1 = _read_field(pre + "_moments", n, M1, S)
    fresh = text.replace('moments = [["totalDegree"], [2]]', 'moments = [["totalDegree"], [3]]').replace(
        'quadPoints = [["tensorizedGrid"], [3]]', 'quadPoints = [["tensorizedGrid"], [5]]').replace(
        "tEnd = 2e-5", "tEnd = 1e3").replace("maxSteps = 2000", "maxSteps = 5")
    pf = tmp_path / "fresh.toml"
    pf.write_text(fresh)
    _, uic, _, _ = native.config_setup(str(pf))
    p2 = tmp_path / "run2.toml"
    p2.write_text(fresh.replace('[general]', '[general]\nrestartFile = "out/run1"'))
    inf, u0, lam0, _ = native.config_setup(str(p2))
    assert (inf["M"], inf["Q"]) == (M2, 25)
    # the oracle's restatement of the start
    m = meshes.naca_ogrid(32, 8)
    b1, bi, b2 = native.Basis(2, [2], 3), native.Basis(2, [2], 5), native.Basis(2, [3], 5)
    c = oracle.closure(oracle.EULER2D, eps=1e-6, safe=1, max_iter=10000)
    uo = u1.copy()
    dirichlet = m.bc == native.BC_DIRICHLET
    assert dirichlet.any()
    uo[dirichlet] = uic[dirichlet, :M1]
    lp = np.zeros_like(uo)
    for j in range(n):
        lp[j, 0] = oracle.DS(c, uo[j, 0])
    st, _ = oracle.solve_closure(c, b1.phi, b1.phiw, uo, lp)
    assert (st == 0).all()
    u2 = oracle.reconstruct(S, b2.phiw, np.ascontiguousarray(oracle.eval_uq(c, bi.phi, lp)))
    l2 = np.zeros((n, M2, S))
    l2[:, :M1] = lp
    c2 = oracle.closure(oracle.EULER2D, eps=1e-6, safe=1, max_iter=1000)
    st, _ = oracle.solve_closure(c2, b2.phi, b2.phiw, u2, l2)
    assert (st == 0).all()
    du = np.abs(u0 - u2).max() / np.abs(u2).max()
    dl = (np.abs(lam0 - l2) / np.abs(l2).max(axis=1, keepdims=True)).max()
    print("start state vs oracle: u", du, "lambda", dl, flush=True)
    assert du < 1e-9 and dl < 1e-7
    csv2 = str(tmp_path / "mv2.csv")
    rep2 = native.run_config(str(p2), csv2)
    assert rep2["steps"] == 5 and rep2["t"] > 2e-5
    co = _oracle_closure(oracle, native.EULER2D, eps=1e-6)
    oprob = oracle.problem(oracle.P_EULER2D, flux=oracle.FLUX_LF, cfl=0.4)
    uo2, lo2 = u2.copy(), l2.copy()
    res = oracle.run(co, oprob, _oracle_mesh(oracle, m), b2.phi, b2.phiw, b2.xi, uo2, lo2, 1e3, 5)
    assert res["rc"] == 0 and res["steps"] == 5
    err = _rel_l2(np.loadtxt(csv2, delimiter=",", ndmin=2), oracle.mean_var(co, b2.phi, b2.phiw, lo2), m.area)
    print("raised-order restart rel L2", err, flush=True)
    assert err.max() < 1e-7


def test_output_files_1d(tmp_path):
    """Burgers: the strip of quadrilaterals (two rows of N+1 points, a tenth of
    the domain high) and the ExpectedValue / Variance text files"""
    text = open(os.path.join(INPUT, "burgers1d_step.toml")).read()
    text = text.replace("numberOfCells = 2000", "numberOfCells = 50").replace("tEnd = 0.11", "tEnd = 0.01")
    p = tmp_path / "b.toml"
    p.write_text(text)
    pre = str(tmp_path / "burgers")
    csv = str(tmp_path / "mv.csv")
    print("case: 1D output files", flush=True)
    native.run_config(str(p), csv, pre)
    mv = np.loadtxt(csv, delimiter=",", ndmin=2)
    pts, cells, types, data = _read_vtk(pre + ".vtk")
    assert pts.shape == (102, 3) and len(cells) == 50 and set(types) == {9}
    assert pts[:51, 1].max() == 0.0 and np.allclose(pts[51:, 1], 0.3)
    np.testing.assert_array_equal(data[("CELL_DATA", "E(u)")][:, 0], mv[:, 0])
    ev = np.loadtxt(pre + "ExpectedValue")
    var = np.loadtxt(pre + "Variance")
    assert ev.shape == var.shape == (50, 2)
    np.testing.assert_allclose(ev[:, 0], np.arange(50) * 3.0 / 50, atol=1e-12)
    np.testing.assert_allclose(ev[:, 1], mv[:, 0], rtol=1e-5)
    np.testing.assert_allclose(var[:, 1], mv[:, 1], rtol=1e-5, atol=1e-12)
