"""Boundary conditions and loads of the benchmark workloads (pyMOTO numbering, ``domain.py:200-251``).

* ``mbb2d``: symmetry roller (x dofs) on x = 0, roller (y dof) at node (nelx, 0), unit downward
  point load on the y dof of node (0, nely) - the MBB half-beam of ``ex_compliance.py``.
* ``cantilever3d``: all dofs of the x = 0 face clamped; a unit total downward (-y) line load on
  the bottom edge (j = 0) of the x = nelx face as consistent nodal forces (1/nelz per node,
  half at both ends).
"""
import numpy as np

from .domain import VoxelDomain


def mbb2d(domain: VoxelDomain):
    nelx, nely = domain.nelx, domain.nely
    left = domain.get_nodenumber(0, np.arange(nely + 1))
    bc = np.concatenate([left * 2, [domain.get_nodenumber(nelx, 0) * 2 + 1]]).astype(np.int64)
    f = np.zeros(2 * domain.nnodes)
    f[domain.get_nodenumber(0, nely) * 2 + 1] = -1.0
    return bc, f


def cantilever3d(domain: VoxelDomain):
    nelx, nely, nelz = domain.nelx, domain.nely, domain.nelz
    jj, kk = np.meshgrid(np.arange(nely + 1), np.arange(nelz + 1), indexing="ij")
    left = domain.get_nodenumber(0, jj, kk).ravel()
    bc = np.sort(np.concatenate([left * 3, left * 3 + 1, left * 3 + 2])).astype(np.int64)
    f = np.zeros(3 * domain.nnodes)
    k = np.arange(nelz + 1)
    w = np.full(nelz + 1, 1.0 / nelz)
    w[[0, -1]] *= 0.5
    f[domain.get_nodenumber(nelx, 0, k) * 3 + 1] = -w
    return bc, f
