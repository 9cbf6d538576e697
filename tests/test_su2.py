"""SU2 mesh ingestion (uqk_su2_mesh, the reader of mesh2d.cpp:68-255) on the
CPU: generated quadrilateral grids against the analytic grid builder,
generated triangle meshes and the reference's shipped NACA mesh against the
invariants every finite-volume mesh satisfies (closed cells, antisymmetric
shared normals, areas adding up), boundary typing by the reference's rule,
and the error paths."""
import os

import numpy as np
import pytest

from uqcreator_amd import mesh as meshes
from uqcreator_amd import native

REF_MESH = "/root/reference/code/input/nacaCoarse.su2"


def _invariants(m):
    n, E = m.nbr.shape
    # closed cells: the scaled outward normals of a cell sum to zero
    assert np.abs(m.normal.sum(1)).max() < 1e-12 * np.abs(m.normal).max()
    # unit normals are unit, scaled ones carry the edge length
    assert np.allclose(np.linalg.norm(m.unit_normal, axis=2), 1.0, atol=1e-14)
    # every interior edge is seen from both sides with opposite normals
    for c in range(n):
        for e in range(E):
            o = m.nbr[c, e]
            if o == n:
                continue
            k = np.flatnonzero(m.nbr[o] == c)
            assert len(k) == 1, (c, e)
            assert np.allclose(m.normal[o, k[0]], -m.normal[c, e], atol=1e-13), (c, e)
    assert (m.area > 0).all()


def test_quad_grid_matches_the_analytic_grid(tmp_path):
    """the same uniform grid as SU2 quadrilaterals and from the grid builder:
    identical neighbours, normals, areas, centres and dt lengths; boundary
    types equal except the four corner cells, which have three marker nodes
    and so are no boundary cells under the reference's rule"""
    p = str(tmp_path / "q.su2")
    nx, ny = 12, 7
    meshes.grid_su2(p, nx, ny, -1.0, 1.0, -0.5, 0.5, triangles=False)
    m = meshes.su2(p, {"left": "dirichlet", "right": "dirichlet", "bottom": "dirichlet", "top": "dirichlet"})
    g = meshes.quad_grid(nx, ny, -1.0, 1.0, -0.5, 0.5, native.BC_DIRICHLET)
    assert m.nbr.shape == g.nbr.shape == (nx * ny, 4)
    np.testing.assert_array_equal(m.nbr, g.nbr)
    np.testing.assert_allclose(m.unit_normal, g.unit_normal, atol=1e-15)
    np.testing.assert_allclose(m.normal, g.normal, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(m.area, g.area, rtol=1e-14)
    np.testing.assert_allclose(m.dx_param, g.dx_param, rtol=1e-14)
    np.testing.assert_allclose(m.center, g.center, atol=1e-15)
    corners = [0, nx - 1, nx * (ny - 1), nx * ny - 1]
    keep = np.setdiff1d(np.arange(nx * ny), corners)
    np.testing.assert_array_equal(m.bc[keep], g.bc[keep])
    assert (m.bc[corners] == native.BC_NONE).all()
    _invariants(m)


def test_triangle_grid_geometry_and_typing(tmp_path):
    p = str(tmp_path / "t.su2")
    nx, ny = 10, 6
    nodes, cells = meshes.grid_su2(p, nx, ny, 0.0, 2.0, 0.0, 1.0, triangles=True)
    m = meshes.su2(p, {"left": "dirichlet", "right": "neumann", "bottom": "noslip", "top": "swwall"})
    assert m.nbr.shape == (2 * nx * ny, 3)
    _invariants(m)
    assert m.area.sum() == pytest.approx(2.0, rel=1e-14)
    np.testing.assert_allclose(m.center, nodes[cells].mean(1), atol=1e-15)
    # dt length: longest edge over area
    L = np.linalg.norm(nodes[cells] - nodes[np.roll(cells, -1, 1)], axis=2).max(1)
    np.testing.assert_allclose(m.dx_param, L / m.area, rtol=1e-14)
    # the reference's boundary cells: exactly two marker nodes
    x, y = nodes[:, 0], nodes[:, 1]
    onb = (np.isclose(x, 0) | np.isclose(x, 2) | np.isclose(y, 0) | np.isclose(y, 1))
    nb = onb[cells].sum(1)
    assert ((m.bc != native.BC_NONE) <= (nb == 2)).all()
    bottom = (nb == 2) & np.isclose(y[cells], 0).sum(1).astype(bool) & ~np.isclose(x[cells], 0).any(1) & \
        ~np.isclose(x[cells], 2).any(1)
    assert (m.bc[bottom] == native.BC_NOSLIP).all() and bottom.sum() > 0
    # the boundary normal is the open edge's (largest neighbour id)
    n = m.n_cells
    for c in np.flatnonzero(m.bc != native.BC_NONE):
        e = int(np.argmax(m.nbr[c]))
        assert m.nbr[c, e] == n
        np.testing.assert_array_equal(m.bnormal[c], m.unit_normal[c, e])


@pytest.mark.skipif(not os.path.exists(REF_MESH), reason="the reference's NACA mesh is not here")
def test_reference_naca_mesh():
    """nacaCoarse.su2 as shipped with the reference: 22,361 triangles, 321 cells
    on the two markers (64 far field, 257 profile)"""
    m = meshes.su2(REF_MESH, {"farfield": "dirichlet", "airfoil": "noslip"})
    assert m.nbr.shape == (22361, 3)
    _invariants(m)
    assert (m.nbr == m.n_cells).sum() == 321
    assert np.bincount(m.bc, minlength=4).tolist() == [22040, 64, 257, 0]


def test_su2_errors(tmp_path):
    p = str(tmp_path / "t.su2")
    meshes.grid_su2(p, 3, 2, 0.0, 1.0, 0.0, 1.0)
    with pytest.raises(RuntimeError, match="no entry in SU2BC"):
        meshes.su2(p, {"left": "dirichlet"})
    with pytest.raises(RuntimeError, match="invalid boundary condition"):
        meshes.su2(p, {"left": "x", "right": "dirichlet", "bottom": "dirichlet", "top": "dirichlet"})
    with pytest.raises(RuntimeError, match="not found"):
        meshes.su2(str(tmp_path / "none.su2"), {})
    nodes = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [2, 0]], float)
    meshes.write_su2(p, nodes, [(0, 1, 2, 3)], {})
    text = open(p).read().replace("NELEM= 1\n9 0 1 2 3 0\n", "NELEM= 2\n9 0 1 2 3 0\n5 1 4 2 1\n")
    open(p, "w").write(text)
    with pytest.raises(RuntimeError, match="mixed"):
        meshes.su2(p, {})


def test_su2_input_file(tmp_path):
    """format = "SU2" in an input file: the mesh file is found next to it and
    the initial state is built on its cell centres"""
    meshes.naca_ogrid_su2(str(tmp_path / "naca_tri.su2"), 32, 8)
    text = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "uqcreator_amd", "input",
                             "euler2d_naca.toml")).read()
    text = text.replace('type = "nacaOGrid"', 'format = "SU2"\nSU2MeshFile = "naca_tri.su2"\n'
                        'SU2BC = [["farfield", "dirichlet"], ["airfoil", "noslip"]]')
    path = tmp_path / "naca_su2.toml"
    path.write_text(text)
    inf, u, lam, center = native.config_setup(str(path))
    m = meshes.su2(str(tmp_path / "naca_tri.su2"), {"farfield": "dirichlet", "airfoil": "noslip"})
    assert inf["edges"] == 3 and u.shape[0] == m.n_cells == 2 * 32 * 8
    np.testing.assert_allclose(center, m.center, atol=0)
