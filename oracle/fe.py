"""Oracle: numbering, element stiffness, CSR assembly, SIMP and sensitivity (test infrastructure only).

Pinning (see tests/test_oracle_fe.py):
* ``dofconn``            - closed form + golden ``dofconn`` arrays written by importing the reference.
* ``element_stiffness``  - one-element tensile closed forms u = FL/(EA), Poisson contraction
                           (``tests/test_assembly.py:101-207`` of the reference), rigid-body null space,
                           symmetry, and the golden K0 from the reference.
* ``csr_pattern``        - brute force COO -> scipy ``tocsr`` pattern, golden indptr/indices (bit-exact).
* ``assemble_values``    - golden CSR data bit-exact, plus dense brute-force sum_e x_e K0.
* ``simp``               - golden xSIMP / dSIMP from the reference's sympy-lambdified expression.
* ``sens_contract``      - finite differences of u^T K(x) v and the golden dC/dx.
* ``element_poisson``    - one-element heat conduction T = Q L / (k A) (2D and 3D), textbook conduction
                           matrices of the unit square / cube, constant null vector, symmetry.
"""
import numpy as np
import scipy.sparse as sps


# --------------------------------------------------------------------------------------------
# numbering (domain.py:200-251, 338-362)
# --------------------------------------------------------------------------------------------
def grid_dim(nelx, nely, nelz):
    return 3 if nelz > 0 else 2


def elem_connectivity(nelx, nely, nelz):
    """(nel, 2**dim) node numbers, local order 0(-,-,-) 1(+,-,-) 2(-,+,-) 3(+,+,-) 4.. at +z."""
    dim = grid_dim(nelx, nely, nelz)
    nz = max(nelz, 1)
    e = np.arange(nelx * nely * nz)
    i, j, k = e % nelx, (e // nelx) % nely, e // (nelx * nely)
    cols = []
    for a in range(2 ** dim):
        di, dj, dk = a & 1, (a >> 1) & 1, (a >> 2) & 1
        cols.append(((k + dk) * (nely + 1) + (j + dj)) * (nelx + 1) + (i + di))
    return np.stack(cols, axis=-1)


def dofconn(nelx, nely, nelz, ndof=None):
    dim = grid_dim(nelx, nely, nelz)
    ndof = dim if ndof is None else ndof
    conn = elem_connectivity(nelx, nely, nelz)
    return (conn[:, :, None] * ndof + np.arange(ndof)[None, None, :]).reshape(conn.shape[0], -1)


def nnodes(nelx, nely, nelz):
    return (nelx + 1) * (nely + 1) * (nelz + 1)


# --------------------------------------------------------------------------------------------
# element stiffness (assembly.py:318-463, domain.py:409-428)
# --------------------------------------------------------------------------------------------
def _shape_fun_der(dim, esize, pos):
    signs = np.array([[1 if (a >> b) & 1 else -1 for b in range(dim)] for a in range(2 ** dim)], dtype=float)
    v = np.prod(esize[:dim])
    dN = np.ones((dim, 2 ** dim)) / v
    for i in range(dim):
        for j in range(dim):
            if i != j:
                dN[i] *= esize[j] / 2 + signs[:, j] * pos[j]
        dN[i] *= signs[:, i]
    return dN


def _B(dN):
    dim, nn = dN.shape
    if dim == 2:
        B = np.zeros((3, 2 * nn))
        for a in range(nn):
            B[:, 2 * a:2 * a + 2] = [[dN[0, a], 0], [0, dN[1, a]], [dN[1, a], dN[0, a]]]
        return B
    B = np.zeros((6, 3 * nn))
    for a in range(nn):
        x, y, z = dN[:, a]
        B[:, 3 * a:3 * a + 3] = [[x, 0, 0], [0, y, 0], [0, 0, z], [0, z, y], [z, 0, x], [y, x, 0]]
    return B


def _D(E, nu, mode):
    mu = E / (2 * (1 + nu))
    lam = (E * nu) / ((1 + nu) * (1 - 2 * nu))
    c1 = 2 * mu + lam
    if mode == "strain":
        return np.array([[c1, lam, 0], [lam, c1, 0], [0, 0, mu]])
    if mode == "stress":
        a = E / (1 - nu * nu)
        return a * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = c1
    D[np.arange(3, 6), np.arange(3, 6)] = mu
    return D


def element_stiffness(dim, esize=(1.0, 1.0, 1.0), E=1.0, nu=0.3, plane="strain"):
    """K0 = sum_gp w B^T D B over 2x2(x2) Gauss points at +-h/(2 sqrt 3); w = prod(h/2) (x h_z in 2D)."""
    esize = np.asarray(esize, dtype=float)
    D = _D(E, nu, "3d" if dim == 3 else plane.lower())
    w = np.prod(esize[:dim] / 2)
    if dim == 2:
        w *= esize[2]
    n = 2 ** dim * dim
    K0 = np.zeros((n, n))
    for a in range(2 ** dim):
        sgn = np.array([1 if (a >> b) & 1 else -1 for b in range(3)], dtype=float)
        pos = sgn * (esize / 2) / np.sqrt(3)
        B = _B(_shape_fun_der(dim, esize, pos))
        K0 += w * B.T @ D @ B
    return K0


def element_poisson(dim, esize=(1.0, 1.0, 1.0), k=1.0):
    """P0 = sum_gp w k dN^T dN over 2x2(x2) Gauss points at +-h/(2 sqrt 3); w = prod(h/2), k times h_z in
    2D (``AssemblePoisson.__init__``, assembly.py:535-560).  Pinned by the reference's one-element heat
    conduction answers (``tests/test_assembly.py:286-340``: T = Q L / (k A)) and the textbook bilinear /
    trilinear conduction matrices of a square / cube."""
    esize = np.asarray(esize, dtype=float)
    w = np.prod(esize[:dim] / 2)
    kk = k * (esize[2] if dim == 2 else 1.0)
    P0 = np.zeros((2 ** dim, 2 ** dim))
    for a in range(2 ** dim):
        sgn = np.array([1 if (a >> b) & 1 else -1 for b in range(3)], dtype=float)
        pos = sgn * (esize / 2) / np.sqrt(3)
        dN = _shape_fun_der(dim, esize, pos)
        P0 += w * kk * dN.T @ dN
    return P0


# --------------------------------------------------------------------------------------------
# CSR assembly with Dirichlet rows/cols (assembly.py:100-296)
# --------------------------------------------------------------------------------------------
class Assembly:
    """Pattern + datamap + np.add.at values, restating ``AssembleGeneral`` for square CSR.

    The pattern is the sorted, de-duplicated union of element DOFs per row (BC rows/cols kept as
    explicit zeros); ``datamap`` sends every element entry to its CSR slot, and BC entries to the
    first BC-diagonal slot, whose value (with every BC diagonal) is then overwritten by ``bcdiagval``
    (``assembly.py:209-230, 267-272``).
    """

    def __init__(self, nelx, nely, nelz, K0, bc=None, bcdiagval=None):
        self.dim = grid_dim(nelx, nely, nelz)
        self.ndof = K0.shape[0] // 2 ** self.dim
        self.n = self.ndof * nnodes(nelx, nely, nelz)
        self.K0 = K0
        self.dc = dofconn(nelx, nely, nelz, self.ndof)
        nel, m = self.dc.shape
        rows = np.repeat(self.dc, m, axis=1).ravel()
        cols = np.tile(self.dc, (1, m)).ravel()
        key = rows.astype(np.int64) * self.n + cols
        ukey, inv = np.unique(key, return_inverse=True)
        self.indices = (ukey % self.n).astype(np.int64)
        urow = ukey // self.n
        self.indptr = np.zeros(self.n + 1, dtype=np.int64)
        np.cumsum(np.bincount(urow, minlength=self.n), out=self.indptr[1:])
        self.datamap = inv.ravel().astype(np.int64)
        self.bc = None if bc is None else np.asarray(bc).ravel()
        self.bcdiagval = bcdiagval
        if self.bc is not None:
            if bcdiagval is None:
                self.bcdiagval = np.max(K0)
            row_of_slot = urow
            isbc = np.zeros(self.n, dtype=bool)
            isbc[self.bc] = True
            bc_slot = isbc[row_of_slot] | isbc[self.indices]
            bc_inds = bc_slot[self.datamap]
            diag_slot = np.flatnonzero((row_of_slot == self.indices) & isbc[row_of_slot])
            # order of bcadd follows self.bc (assembly.py:221-228)
            slot_of_row = {int(r): int(s) for r, s in zip(row_of_slot[diag_slot], diag_slot)}
            self.bcadd = np.array([slot_of_row[int(b)] for b in self.bc], dtype=np.int64)
            self.datamap[bc_inds] = self.bcadd[0]

    def values(self, x):
        """CSR data: np.add.at of x_e*K0_ij in element-major order, then BC diagonal overwrite."""
        scaled = (self.K0.ravel()[None, :] * x.ravel()[:, None]).ravel()
        data = np.zeros(self.indices.size, dtype=np.result_type(x, self.K0))
        np.add.at(data, self.datamap, scaled)
        if self.bc is not None:
            data[self.bcadd] = self.bcdiagval
        return data

    def matrix(self, x):
        return sps.csr_matrix((self.values(x), self.indices, self.indptr), shape=(self.n, self.n))

    def sens_contract(self, u_list, v_list):
        """dx_e = sum_k u_k[dofs_e]^T K0 v_k[dofs_e] with BC entries zeroed (assembly.py:298-315)."""
        dx = np.zeros(self.dc.shape[0])
        for u, v in zip(u_list, v_list):
            u = np.array(u, dtype=float, copy=True)
            v = np.array(v, dtype=float, copy=True)
            if self.bc is not None:
                u[self.bc] = 0.0
                v[self.bc] = 0.0
            dx += np.einsum("Ai,ij,Aj->A", u[self.dc], self.K0, v[self.dc])
        return dx


# --------------------------------------------------------------------------------------------
# SIMP  y = xmin + (1 - xmin) x^p  (generic.py:94-140 with the expression of ex_compliance.py:89)
# --------------------------------------------------------------------------------------------
def simp(x, xmin=1e-9, p=3.0):
    return xmin + (1.0 - xmin) * x ** p


def dsimp(x, xmin=1e-9, p=3.0):
    return p * (1.0 - xmin) * x ** (p - 1)
