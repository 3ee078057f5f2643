"""Oracle-side definitions of the benchmark problems (test infrastructure only).

Boundary conditions and loads of the workloads, written with the reference's numbering
(``domain.py:200-251``).  ``tests/test_problems.py`` checks that the package's own
``pymoto_amd.problems`` builds the very same index sets and load vectors.

* MBB beam (2D, Q4): x-DOF of every node on x=0 (symmetry roller), y-DOF of node (nelx, 0) (roller),
  unit downward point load on the y-DOF of node (0, nely).
* Cantilever (3D, hex8): every DOF of every node on x=0 clamped; a unit total downward (-y) line load
  along the bottom edge (j=0) of the x=nelx face, as consistent nodal forces (1/nelz per node,
  halved at the two end nodes).
"""
import numpy as np


def node(nelx, nely, i, j, k=0):
    return (k * (nely + 1) + j) * (nelx + 1) + i


def mbb2d(nelx, nely):
    nn = (nelx + 1) * (nely + 1)
    left = node(nelx, nely, 0, np.arange(nely + 1))
    bc = np.concatenate([left * 2, [node(nelx, nely, nelx, 0) * 2 + 1]]).astype(np.int64)
    f = np.zeros(2 * nn)
    f[node(nelx, nely, 0, nely) * 2 + 1] = -1.0
    return bc, f


def cantilever3d(nelx, nely, nelz):
    nn = (nelx + 1) * (nely + 1) * (nelz + 1)
    jj, kk = np.meshgrid(np.arange(nely + 1), np.arange(nelz + 1), indexing="ij")
    left = node(nelx, nely, 0, jj, kk).ravel()
    bc = np.sort(np.concatenate([left * 3, left * 3 + 1, left * 3 + 2])).astype(np.int64)
    f = np.zeros(3 * nn)
    kload = np.arange(nelz + 1)
    w = np.full(nelz + 1, 1.0 / nelz)
    w[0] *= 0.5
    w[-1] *= 0.5
    f[node(nelx, nely, nelx, 0, kload) * 3 + 1] = -w
    return bc, f


def face_point_loads(nelx, nely, nelz, nloads, seed):
    """k unit downward (-y) point loads at seeded nodes (j, k) of the x = nelx face, one per column of
    F (ndof, k): the load cases of the batched robust workload."""
    rng = np.random.default_rng(seed)
    nn = (nelx + 1) * (nely + 1) * (nelz + 1)
    F = np.zeros((3 * nn, nloads))
    for c in range(nloads):
        j, k = int(rng.integers(0, nely + 1)), int(rng.integers(0, nelz + 1))
        F[3 * node(nelx, nely, nelx, j, k) + 1, c] = -1.0
    return F
