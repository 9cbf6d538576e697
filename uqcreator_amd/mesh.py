"""Mesh arrays for the flux kernels: neighbours, normals, areas, boundary types.

Array conventions (shared with include/uqk.h's uqk_mesh):
  nbr[j, e]            neighbour across edge e of cell j; n_cells marks a boundary
                       edge (mesh2d.cpp:194)
  unit_normal[j, e]    outward unit normal; normal[j, e] = length * unit normal
  area[j]              cell measure
  dx_param[j]          maxEdge / area, the argument of Problem::ComputeDt
                       (momentsolver.cpp:130)
  bc[j]                boundary type; bnormal[j] the boundary unit normal
  center[j]            the point the initial condition is evaluated at
"""
import numpy as np

from . import native


class MeshArrays:
    def __init__(self, nbr, unit_normal, normal, area, dx_param, bc, bnormal, center):
        self.nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.unit_normal = np.ascontiguousarray(unit_normal, dtype=np.float64)
        self.normal = np.ascontiguousarray(normal, dtype=np.float64)
        self.area = np.ascontiguousarray(area, dtype=np.float64)
        self.dx_param = np.ascontiguousarray(dx_param, dtype=np.float64)
        self.bc = np.ascontiguousarray(bc, dtype=np.int32)
        self.bnormal = np.ascontiguousarray(bnormal, dtype=np.float64)
        self.center = np.ascontiguousarray(center, dtype=np.float64)

    @property
    def n_cells(self):
        return self.nbr.shape[0]


def su2(path, bcs):
    """An SU2 mesh file (mesh2d.cpp:68-255) read by the library's C++ reader
    (uqk_su2_mesh); bcs maps marker tags to dirichlet | noslip | swwall |
    neumann | periodic."""
    import ctypes as C

    spec = ";".join("%s=%s" % kv for kv in dict(bcs).items()).encode()
    err = C.create_string_buffer(512)
    E = C.c_int32()
    lib = native.lib()
    n = lib.uqk_su2_mesh(path.encode(), spec, C.byref(E), *([None] * 8), err, 512)
    if n < 0:
        raise RuntimeError(err.value.decode())
    E = E.value
    nbr = np.zeros((n, E), np.int32)
    un = np.zeros((n, E, 2))
    sn = np.zeros((n, E, 2))
    area = np.zeros(n)
    dx = np.zeros(n)
    bc = np.zeros(n, np.int32)
    bn = np.zeros((n, 2))
    center = np.zeros((n, 2))
    p = native._p
    n2 = lib.uqk_su2_mesh(path.encode(), spec, None, p(nbr), p(un), p(sn), p(area), p(dx), p(bc), p(bn), p(center),
                          err, 512)
    if n2 != n:
        raise RuntimeError(err.value.decode())
    return MeshArrays(nbr, un, sn, area, dx, bc, bn, center)


def write_su2(path, nodes, cells, markers):
    """Writes an SU2 mesh: nodes (P, 2), cells (N, 3) triangles or (N, 4)
    quadrilaterals of node indices, markers {tag: (K, 2) node pairs}."""
    cells = np.asarray(cells)
    etype = {3: 5, 4: 9}[cells.shape[1]]
    with open(path, "w") as f:
        f.write("NDIME= 2\nNELEM= %d\n" % len(cells))
        for i, c in enumerate(cells):
            f.write("%d %s %d\n" % (etype, " ".join(str(int(v)) for v in c), i))
        f.write("NPOIN= %d\n" % len(nodes))
        for i, (x, y) in enumerate(nodes):
            f.write("%.17g %.17g %d\n" % (x, y, i))
        f.write("NMARK= %d\n" % len(markers))
        for tag, pairs in markers.items():
            f.write("MARKER_TAG= %s\nMARKER_ELEMS= %d\n" % (tag, len(pairs)))
            for a, b in pairs:
                f.write("3 %d %d\n" % (a, b))


def grid_su2(path, nx, ny, x0, x1, y0, y1, triangles=True, tags=("left", "right", "bottom", "top")):
    """A uniform nx x ny grid of [x0,x1] x [y0,y1] written as SU2 (each
    square split into two triangles, or kept as a quadrilateral), with one
    marker per side.  Returns (nodes, cells)."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel()], 1)
    nid = lambda i, j: i * (ny + 1) + j
    cells = []
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if triangles:
                cells += [(a, b, c), (a, c, d)] if (i + j) % 2 == 0 else [(a, b, d), (b, c, d)]
            else:
                cells.append((a, b, c, d))
    markers = {tags[0]: [(nid(0, j), nid(0, j + 1)) for j in range(ny)],
               tags[1]: [(nid(nx, j), nid(nx, j + 1)) for j in range(ny)],
               tags[2]: [(nid(i, 0), nid(i + 1, 0)) for i in range(nx)],
               tags[3]: [(nid(i, ny), nid(i + 1, ny)) for i in range(nx)]}
    write_su2(path, nodes, cells, markers)
    return nodes, np.asarray(cells)


def mesh1d(n_cells, a, b):
    """Mesh1D (mesh1d.cpp:5-67, line.cpp:3-44): nodes a + i (b-a)/N, cells 0 and
    N-1 Dirichlet, the cell 'centre' is its left node (line.cpp:17) and
    maxEdge = area^2 (line.cpp:18-19) so dx_param = area."""
    N = n_cells
    nodes = a + np.arange(N + 1) * (b - a) / N
    area = np.abs(nodes[1:] - nodes[:-1])
    nbr = np.stack([np.arange(N) - 1, np.arange(N) + 1], 1)
    bc = np.zeros(N, dtype=np.int32)
    bc[0] = bc[-1] = native.BC_DIRICHLET
    nbr[0] = N
    nbr[-1] = N
    un = np.zeros((N, 2, 1))
    un[:, 0, 0] = -1.0
    un[:, 1, 0] = 1.0
    return MeshArrays(nbr, un, un.copy(), area, area ** 2 / area, bc, np.zeros((N, 1)), nodes[:-1, None])


def _quad_mesh(X, Y, periodic_i, bc_j0, bc_j1, bc_i0=None, bc_i1=None):
    """Quadrilateral mesh from node arrays X, Y of shape (ni+1, nj+1) (or
    (ni, nj+1) with periodic_i).  Cell (i, j) has id j*ni + i and nodes
    (i,j), (i+1,j), (i+1,j+1), (i,j+1); edge e joins node e and node e+1, so the
    edges face j-, i+, j+, i-.  Outward normals are fixed by the centroid test
    of Quadrangle::getOutwardNormal (quadrangle.cpp:72-83); the area is the
    shoelace formula and dx_param uses the longest edge (the reference's
    Quadrangle never updates maxEdge, quadrangle.cpp:58,69)."""
    ni = X.shape[0] if periodic_i else X.shape[0] - 1
    nj = X.shape[1] - 1
    N = ni * nj
    ii, jj = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
    ii = ii.T.ravel()
    jj = jj.T.ravel()
    ip = (ii + 1) % X.shape[0] if periodic_i else ii + 1
    px = np.stack([X[ii, jj], X[ip, jj], X[ip, jj + 1], X[ii, jj + 1]], 1)
    py = np.stack([Y[ii, jj], Y[ip, jj], Y[ip, jj + 1], Y[ii, jj + 1]], 1)
    cx, cy = px.mean(1), py.mean(1)
    area = 0.5 * np.abs(np.sum(px * np.roll(py, -1, 1) - np.roll(px, -1, 1) * py, 1))
    dx = np.roll(px, -1, 1) - px
    dy = np.roll(py, -1, 1) - py
    length = np.hypot(dx, dy)
    nx, ny = dy.copy(), -dx.copy()
    flip = nx * (cx[:, None] - px) + ny * (cy[:, None] - py) > 0
    nx[flip] *= -1
    ny[flip] *= -1
    nrm = np.hypot(nx, ny)
    un = np.stack([nx / nrm, ny / nrm], -1)
    sn = un * length[..., None]
    cid = lambda i, j: j * ni + i
    nbr = np.empty((N, 4), dtype=np.int64)
    nbr[:, 0] = np.where(jj > 0, cid(ii, jj - 1), N)
    nbr[:, 2] = np.where(jj < nj - 1, cid(ii, jj + 1), N)
    if periodic_i:
        nbr[:, 1] = cid((ii + 1) % ni, jj)
        nbr[:, 3] = cid((ii - 1) % ni, jj)
    else:
        nbr[:, 1] = np.where(ii < ni - 1, cid(ii + 1, jj), N)
        nbr[:, 3] = np.where(ii > 0, cid(ii - 1, jj), N)
    bc = np.zeros(N, dtype=np.int32)
    bc[jj == 0] = bc_j0
    bc[jj == nj - 1] = bc_j1
    if not periodic_i:
        bc[ii == 0] = bc_i0
        bc[ii == ni - 1] = bc_i1
    # boundary normal: the edge with the largest neighbour id (cell.cpp:57-68)
    first_max = np.argmax(nbr, axis=1)
    bnormal = un[np.arange(N), first_max]
    return MeshArrays(nbr, un, sn, area, length.max(1) / area, bc, bnormal, np.stack([cx, cy], 1))


def quad_grid(nx, ny, x0, x1, y0, y1, bc_boundary):
    """Uniform Cartesian quads on [x0,x1] x [y0,y1]; every boundary cell gets bc_boundary."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return _quad_mesh(X, Y, False, bc_boundary, bc_boundary, bc_boundary, bc_boundary)


def naca0012(x):
    """Half thickness of NACA0012 with the closed trailing edge (-0.1036 x^4)."""
    return 0.6 * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)


def naca_ogrid(ni=512, nj=128, radius=20.0, first_height=2e-3):
    """Structured O-grid around an analytic NACA0012 of unit chord: ni cells
    around the profile (cosine clustered at both edges, starting at the
    trailing edge and running along the lower side first), nj cells out to a
    circle of `radius` chords centred at mid chord, geometric stretching from
    `first_height`.  The profile ring is a slip wall (the reference's 'noslip'
    ghost, expliciteuler.cpp:25-43); the outer ring is Dirichlet far field."""
    X, Y = naca_ogrid_nodes(ni, nj, radius, first_height)
    return _quad_mesh(X, Y, True, native.BC_NOSLIP, native.BC_DIRICHLET)


def naca_ogrid_su2(path, ni=64, nj=16, radius=20.0, first_height=2e-3):
    """The nodes of naca_ogrid as an SU2 triangle mesh (each quadrilateral
    split along alternating diagonals), markers 'airfoil' (profile ring) and
    'farfield' (outer ring).  Returns (nodes, cells)."""
    X, Y = naca_ogrid_nodes(ni, nj, radius, first_height)
    nodes = np.stack([X.ravel(), Y.ravel()], 1)  # node (i, j) at i*(nj+1) + j
    nid = lambda i, j: (i % ni) * (nj + 1) + j
    cells = []
    for j in range(nj):
        for i in range(ni):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            cells += [(a, b, c), (a, c, d)] if (i + j) % 2 == 0 else [(a, b, d), (b, c, d)]
    markers = {"airfoil": [(nid(i, 0), nid(i + 1, 0)) for i in range(ni)],
               "farfield": [(nid(i, nj), nid(i + 1, nj)) for i in range(ni)]}
    write_su2(path, nodes, cells, markers)
    return nodes, np.asarray(cells)


def naca_ogrid_nodes(ni=512, nj=128, radius=20.0, first_height=2e-3):
    """Node coordinates X, Y (ni x nj+1, periodic in i) of naca_ogrid."""
    half = ni // 2
    th = np.linspace(0.0, np.pi, half + 1)
    # half cosine, half uniform spacing: refined at both edges without
    # sliver cells at the trailing edge
    xs = 0.5 * (0.5 * (1 + np.cos(th)) + (1 - th / np.pi))  # 1 -> 0
    lower_x, lower_y = xs, -naca0012(xs)    # TE -> LE along the lower side
    xu = xs[::-1][1:-1]                      # LE -> TE (exclusive)
    upper_x, upper_y = xu, naca0012(xu)
    sx = np.concatenate([lower_x, upper_x])
    sy = np.concatenate([lower_y, upper_y])
    assert len(sx) == ni
    ang = np.arctan2(sy, sx - 0.5)
    fx = 0.5 + radius * np.cos(ang)
    fy = radius * np.sin(ang)
    # geometric stretching s_0 = 0, s_nj = 1
    dist = np.hypot(fx - sx, fy - sy).mean()
    h0 = first_height / dist
    lo, hi = 1.0 + 1e-9, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        tot = h0 * (r ** nj - 1) / (r - 1)
        lo, hi = (r, hi) if tot < 1.0 else (lo, r)
    s = np.concatenate([[0.0], h0 * (r ** np.arange(1, nj + 1) - 1) / (r - 1)])
    s /= s[-1]
    X = sx[:, None] + (fx - sx)[:, None] * s[None, :]
    Y = sy[:, None] + (fy - sy)[:, None] * s[None, :]
    return X, Y
