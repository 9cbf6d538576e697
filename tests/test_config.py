"""Input files (the reference's TOML schema, settings.cpp:22-323): the host
layer parses them, builds the mesh and the initial moments exactly as the
Python workloads do, and rejects what the reference rejects.  No GPU work."""
import os

import numpy as np
import pytest

from uqcreator_amd import mesh as meshes
from uqcreator_amd import native, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INPUT = os.path.join(ROOT, "uqcreator_amd", "input")


def _write(tmp_path, text, name="case.toml"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_burgers_input_matches_workload():
    inf, u, lam, center = native.config_setup(os.path.join(INPUT, "burgers1d_step.toml"))
    assert (inf["closure"], inf["S"], inf["M"], inf["Q"], inf["d"], inf["edges"]) == (0, 1, 6, 12, 1, 2)
    b, m, u0, lam0 = workloads.burgers_c1(2000)
    np.testing.assert_array_equal(center, m.center)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(lam, lam0)


def test_euler2d_input_matches_workload_and_ds(tmp_path, oracle):
    text = open(os.path.join(INPUT, "euler2d_naca.toml")).read()
    text = text.replace("ni = 512", "ni = 64").replace("nj = 128", "nj = 16")
    inf, u, lam, center = native.config_setup(_write(tmp_path, text))
    assert (inf["closure"], inf["S"], inf["M"], inf["Q"], inf["d"], inf["edges"], inf["precision"]) == \
        (native.EULER2D, 4, 10, 25, 2, 4, native.FP64)
    b, m, u0 = workloads.euler2d_c3(64, 16)
    np.testing.assert_allclose(center, m.center, rtol=0, atol=1e-12)
    np.testing.assert_allclose(u, u0, rtol=1e-13)
    c = oracle.closure(oracle.EULER2D)
    for j in (0, 100, 1023):
        np.testing.assert_allclose(lam[j, 0], oracle.DS(c, u[j, 0]), rtol=1e-12)
        assert not lam[j, 1:].any()
    # the DS start is exact for the constant moment: U(lambda_0) = u_0
    np.testing.assert_allclose(oracle.U(c, lam[0, 0]), u[0, 0], rtol=1e-12)


def test_full_size_naca_inputs():
    inf, u, lam, center = native.config_setup(os.path.join(INPUT, "euler2d_naca_mixed.toml"))
    assert inf["precision"] == native.MIXED and u.shape == (512 * 128, 10, 4)
    inf64, u64, _, _ = native.config_setup(os.path.join(INPUT, "euler2d_naca.toml"))
    assert inf64["precision"] == native.FP64
    np.testing.assert_array_equal(u, u64)


def test_mesh_builders_match_python(tmp_path):
    text = open(os.path.join(INPUT, "shallowwater2d_dambreak.toml")).read()
    text = text.replace("nx = 1024", "nx = 24").replace("ny = 1024", "ny = 20")
    inf, u, lam, center = native.config_setup(_write(tmp_path, text))
    assert (inf["closure"], inf["S"], inf["M"], inf["Q"], inf["d"]) == (native.SHALLOW_WATER2D_BARRIER, 3, 35, 125, 3)
    m = meshes.quad_grid(24, 20, -1, 1, -1, 1, native.BC_DIRICHLET)
    np.testing.assert_allclose(center, m.center, atol=1e-14)
    b = native.Basis(3, [4], 5)
    u0 = workloads.project_ic(workloads.sw2d_dambreak_ic(), m.center, b)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-13)
    # barrier DS: lambda_0 = (g h - barrier/h - |v|^2/2, v)
    h = u[:, 0, 0]
    np.testing.assert_allclose(lam[:, 0, 0], 9.81 * h - 0.01 / h, rtol=1e-13)


@pytest.mark.parametrize("edit,msg", [
    (("problem = \"Burgers1D\"", "problem = \"Heat\""), "problem"),
    (("closure = \"BoundedBarrier\"", "closure = \"LogSin\""), "closure"),
    (("quadPoints = [[\"tensorizedGrid\"], [12]]", "quadPoints = [[\"sparseGrid\"], [12]]"), "sparse"),
    (("CFL = 0.5", ""), "CFL"),
    (("distribution = [[\"Legendre\"], [0.2]]", "distribution = [[\"Hermite\"], [0.2]]"), "Legendre"),
    (("numberOfCells = 2000", "numberOfCells = 1"), "mesh"),
    (("tEnd = 0.11", "tEnd = 0.11 0.2"), "toml line"),
])
def test_input_errors_are_reported(tmp_path, edit, msg):
    text = open(os.path.join(INPUT, "burgers1d_step.toml")).read()
    assert edit[0] in text
    with pytest.raises(RuntimeError, match=msg):
        native.config_setup(_write(tmp_path, text.replace(edit[0], edit[1])))


def test_missing_file_is_reported():
    with pytest.raises(RuntimeError, match="cannot open"):
        native.config_setup("/nonexistent/input.toml")


def test_toml_subset(tmp_path):
    """comments, underscores in numbers, multi-line nested arrays, single quotes"""
    text = open(os.path.join(INPUT, "burgers1d_step.toml")).read()
    text = text.replace("numberOfCells = 2000", "numberOfCells = 2_000  # cells")
    text = text.replace("distribution = [[\"Legendre\"], [0.2]]", "distribution = [\n  ['Legendre'],\n  [0.2],\n]")
    inf, u, lam, center = native.config_setup(_write(tmp_path, text))
    assert center.shape == (2000, 1)
