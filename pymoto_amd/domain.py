"""Structured voxel grid (Q4 in 2D, hex8 in 3D): numbering, connectivity and shape functions.

Mirrors ``pymoto/common/domain.py`` (``VoxelDomain``):

* element number ``e = (k*nely + j)*nelx + i``      (``domain.py:200-211``)
* node number    ``n = (k*(nely+1) + j)*(nelx+1) + i`` (``domain.py:213-224``)
* DOF number     ``n*ndof + d``                      (``domain.py:226-251``)
* local node order 0(-,-,-) 1(+,-,-) 2(-,+,-) 3(+,+,-), 4..7 the same at +z (``domain.py:131-142``)

Only setup-time host arrays live here (``conn``, ``elements``, ``nodes``); the device kernels
recompute every index in closed form from (i, j, k) and never read these arrays.
"""
import warnings
from typing import Iterable, Union

import numpy as np


class VoxelDomain:
    """Structured 2D/3D voxel domain (``domain.py:57-164``)."""

    def __init__(self, nelx: int, nely: int, nelz: int = 0, unitx: float = 1.0, unity: float = 1.0,
                 unitz: float = 1.0):
        self.nelx, self.nely, self.nelz = int(nelx), int(nely or 0), int(nelz or 0)
        self.dim = 1 if (self.nelz == 0 and self.nely == 0) else (2 if self.nelz == 0 else 3)
        self.origin = np.array([0.0, 0.0, 0.0])
        self.unitx, self.unity, self.unitz = float(unitx), float(unity), float(unitz)
        assert np.prod(self.element_size[: self.dim]) > 0.0, "Element volume needs to be positive"

        self.nel = self.nelx * self.nely * max(self.nelz, 1)
        self.nnodes = (self.nelx + 1) * (self.nely + 1) * (self.nelz + 1)
        self.elemnodes = 2 ** self.dim

        # Local node sign pattern: bit 0 -> x, bit 1 -> y, bit 2 -> z (domain.py:131-142)
        self.node_numbering = [[0, 0, 0] for _ in range(self.elemnodes)]
        for a in range(self.elemnodes):
            self.node_numbering[a] = [1 if (a >> b) & 1 else -1 for b in range(3)]
            if self.dim < 3:
                self.node_numbering[a][2] = -1
            if self.dim < 2:
                self.node_numbering[a][1] = -1

        eli, elj, elk = np.meshgrid(np.arange(self.nelx), np.arange(max(self.nely, 1)),
                                    np.arange(max(self.nelz, 1)), indexing="ij")
        self.elements = self.get_elemnumber(eli, elj, elk)
        ndi, ndj, ndk = np.meshgrid(np.arange(self.nelx + 1), np.arange(self.nely + 1),
                                    np.arange(self.nelz + 1), indexing="ij")
        self.nodes = self.get_nodenumber(ndi, ndj, ndk)  # (nelx+1, nely+1, nelz+1), also in 2D
        self._conn = None

    # ----- sizes -------------------------------------------------------------------------------
    @property
    def element_size(self):
        return np.array([self.unitx, self.unity, self.unitz])

    @property
    def domain_size(self):
        return np.array([self.nelx * self.unitx, self.nely * self.unity, self.nelz * self.unitz])[: self.dim]

    @property
    def size(self):
        return np.array([self.nelx, self.nely, self.nelz])[: self.dim]

    @property
    def conn(self):
        """Element-to-node connectivity ``(nel, 2**dim)`` (``domain.py:150-152``), built lazily."""
        if self._conn is None:
            e = np.arange(self.nel)
            i, j, k = e % self.nelx, (e // self.nelx) % max(self.nely, 1), e // (self.nelx * max(self.nely, 1))
            self._conn = self.get_elemconnectivity(i, j, k)
        return self._conn

    # ----- numbering ---------------------------------------------------------------------------
    def get_elemnumber(self, eli, elj, elk=0):
        return (elk * self.nely + elj) * self.nelx + eli

    def get_nodenumber(self, nodi, nodj, nodk=0):
        return (nodk * (self.nely + 1) + nodj) * (self.nelx + 1) + nodi

    def get_dofnumber(self, nod_idx, dof_idx=None, ndof: int = None):
        if not isinstance(nod_idx, (int, np.integer)):
            nod_idx = np.asarray(nod_idx)
        ndof = self.dim if ndof is None else ndof
        if dof_idx is None:
            dof_idx = np.arange(ndof)
        if not isinstance(dof_idx, (int, np.integer)):
            dof_idx = np.asarray(dof_idx)
        if np.ndim(dof_idx) == 0 or np.ndim(nod_idx) == 0:
            return nod_idx * ndof + dof_idx
        n1 = np.expand_dims(nod_idx, axis=tuple(-(np.arange(np.ndim(dof_idx)) + 1)))
        d1 = np.expand_dims(dof_idx, axis=tuple(np.arange(np.ndim(nod_idx))))
        return n1 * ndof + d1

    def get_node_indices(self, nod_idx=None):
        if nod_idx is None:
            nod_idx = np.arange(self.nnodes)
        nodi = nod_idx % (self.nelx + 1)
        nodj = (nod_idx // (self.nelx + 1)) % (self.nely + 1)
        if self.dim == 2:
            return np.stack([nodi, nodj], axis=0)
        nodk = nod_idx // ((self.nelx + 1) * (self.nely + 1))
        return np.stack([nodi, nodj, nodk], axis=0)

    def get_node_position(self, nod_idx=None):
        ijk = self.get_node_indices(nod_idx)
        return (self.origin[: self.dim] + self.element_size[: self.dim] * ijk.T).T

    def get_element_indices(self, el_idx=None):
        if el_idx is None:
            el_idx = np.arange(self.nel)
        eli = el_idx % self.nelx
        elj = (el_idx // self.nelx) % self.nely
        if self.dim == 2:
            return np.stack([eli, elj], axis=0)
        elk = el_idx // (self.nelx * self.nely)
        return np.stack([eli, elj, elk], axis=0)

    def get_elemconnectivity(self, i, j, k=0):
        nods = [self.get_nodenumber(i + max(n[0], 0), j + max(n[1], 0), k + max(n[2], 0))
                for n in self.node_numbering]
        return np.stack(nods, axis=-1)

    def get_dofconnectivity(self, ndof: int):
        return np.reshape(self.get_dofnumber(self.conn, ndof=ndof), (self.conn.shape[0], -1))

    # ----- shape functions (domain.py:364-428) -------------------------------------------------
    def eval_shape_fun(self, pos: np.ndarray):
        v = np.prod(self.element_size[: self.dim])
        shapefn = np.ones(self.elemnodes) / v
        for i in range(self.dim):
            shapefn *= np.array([self.element_size[i] / 2 + n[i] * pos[i] for n in self.node_numbering])
        return shapefn

    def eval_shape_fun_der(self, pos: np.ndarray):
        v = np.prod(self.element_size[: self.dim])
        dN = np.ones((self.dim, self.elemnodes)) / v
        for i in range(self.dim):
            for j in range(self.dim):
                if i != j:
                    dN[i, :] *= np.array([self.element_size[j] / 2 + n[j] * pos[j] for n in self.node_numbering])
            dN[i, :] *= np.array([n[i] for n in self.node_numbering])
        return dN

    def get_rigid_body_modes(self):
        nrbm = self.dim * (self.dim + 1) // 2
        rbm = np.zeros((self.nnodes * self.dim, nrbm))
        for i in range(self.dim):
            rbm[i::self.dim, i] = 1.0
        c = self.get_node_position()
        if self.dim == 3:
            rbm[1::3, -3] = c[2]
            rbm[2::3, -3] = -c[1]
            rbm[0::3, -2] = c[2]
            rbm[2::3, -2] = -c[0]
        rbm[0::self.dim, -1] = -c[1]
        rbm[1::self.dim, -1] = c[0]
        return rbm

    def write_to_vti(self, vectors: dict, filename="out.vti", scale=1.0):
        """Write cell/point vectors (numpy or device tensors) to a Paraview VTI file (``domain.py:452-604``)."""
        from .io import write_to_vti
        write_to_vti(self, vectors, filename, scale)

    def coarsen(self):
        """The 2x coarser grid used by geometric multigrid (``iterative.py:163-165``)."""
        return VoxelDomain(self.nelx // 2, self.nely // 2, self.nelz // 2,
                           self.unitx * 2, self.unity * 2, self.unitz * 2)

    def __repr__(self):
        dims = f"{self.nelx}x{self.nely}" + (f"x{self.nelz}" if self.dim == 3 else "")
        return f"VoxelDomain({dims}, nel={self.nel}, nnodes={self.nnodes})"
