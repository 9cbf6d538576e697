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
