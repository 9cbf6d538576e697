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


def test_shallow_water_input_short_run(oracle, tmp_path):
    """The dam-break input (barrier closure, n = 105, 5^3 points, bed friction
    on the third uncertain dimension) on a 16 x 16 grid, 4 steps, against the
    oracle's time loop."""
    text = open(os.path.join(INPUT, "shallowwater2d_dambreak.toml")).read()
    for a, b in [("nx = 1024", "nx = 16"), ("ny = 1024", "ny = 16"), ("maxSteps = 500", "maxSteps = 4"),
                 ('precision = "mixed"', 'precision = "fp64"')]:
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
    print("report", rep, "rel L2 E/Var", err, flush=True)
    assert rep["steps"] == 4
    assert err.max() < 1e-8


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
