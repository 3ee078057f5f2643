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


def face_point_loads(domain: VoxelDomain, nloads: int, seed: int):
    """``nloads`` unit downward (-y) point loads at seeded nodes (j, k) of the x = nelx face, one per
    column of F (ndof, nloads): the load cases of the batched robust workload."""
    rng = np.random.default_rng(seed)
    F = np.zeros((3 * domain.nnodes, nloads))
    for c in range(nloads):
        j, k = int(rng.integers(0, domain.nely + 1)), int(rng.integers(0, domain.nelz + 1))
        F[3 * domain.get_nodenumber(domain.nelx, j, k) + 1, c] = -1.0
    return F


def uneven_batch_shapes(nproblems: int, seed: int, reduced: bool = False):
    """Grid shapes of the uneven batch of independent problems, drawn from ``seed``: half-and-half Q4 beams
    (nx x nx/2) and hex8 cantilevers (nx x nx/2 x nx/2), from 64x32 Q4 to 96x48x48 hex8 (``reduced``: from
    16x8 Q4 to 24x12x12 hex8, the sizes the parity tests solve directly on the host)."""
    rng = np.random.default_rng(seed)
    q4 = [16, 24, 32, 48, 64] if reduced else [64, 96, 128, 160, 192, 256]
    hx = [8, 12, 16, 20, 24] if reduced else [32, 48, 64, 80, 96]
    out = []
    for _ in range(nproblems):
        if rng.random() < 0.5:
            nx = int(rng.choice(q4))
            out.append((nx, nx // 2, 0))
        else:
            nx = int(rng.choice(hx))
            out.append((nx, nx // 2, nx // 2))
    return out
