"""Stiffness assembly on the GPU: ``AssembleStiffness`` / ``AssembleGeneral``.

Mirrors ``pymoto/modules/assembly.py``:

* ``AssembleStiffness(domain, *args, e_modulus=1.0, poisson_ratio=0.3, plane='strain', **kwargs)``
  integrates the element matrix K0 = sum_gp w B^T D B with 2x2(x2) Gauss points
  (``assembly.py:418-463``, ``get_B`` 318-370, ``get_D`` 373-404).
* ``AssembleGeneral(domain, element_matrix, bc=None, bcdiagval=None, matrix_type=csr_matrix,
  add_constant=None, reuse_sparsity=True)`` maps densities x to K(x) = sum_e x_e K0 with Dirichlet
  rows/cols zeroed and ``bcdiagval`` (default max K0) on their diagonal (``assembly.py:38-296``).

The output is a :class:`StiffnessMatrix`: the operator stays matrix-free on the device (x and K0
are all it stores) and is applied by the HIP node-centric kernel.  ``tocsr()`` materialises the
reference's exact CSR arrays on the device when a caller needs them.  The adjoint
``_sensitivity`` contracts a :class:`DyadicMatrix` with K0 per element on the GPU
(``assembly.py:298-315``).
"""
import numpy as np
import torch

from . import _lib
from .core import Module
from .domain import VoxelDomain
from .dyadic import DyadicMatrix


def get_B(dN_dx):
    """Strain-displacement matrix, Voigt order [xx, yy, zz, yz, zx, xy] in 3D (``assembly.py:318-370``)."""
    n_dim, n_shape = dN_dx.shape
    B = np.zeros((n_dim * (n_dim + 1) // 2, n_shape * n_dim))
    for a in range(n_shape):
        if n_dim == 2:
            dx, dy = dN_dx[:, a]
            B[:, 2 * a:2 * a + 2] = [[dx, 0.0], [0.0, dy], [dy, dx]]
        elif n_dim == 3:
            dx, dy, dz = dN_dx[:, a]
            B[:, 3 * a:3 * a + 3] = [[dx, 0, 0], [0, dy, 0], [0, 0, dz], [0, dz, dy], [dz, 0, dx], [dy, dx, 0]]
        else:
            B[0, a] = dN_dx[0, a]
    return B


def get_D(E: float, nu: float, mode: str = "strain"):
    """Isotropic constitutive matrix: plane strain, plane stress or 3D (``assembly.py:373-404``)."""
    mu = E / (2 * (1 + nu))
    lam = (E * nu) / ((1 + nu) * (1 - 2 * nu))
    c1 = 2 * mu + lam
    mode = mode.lower()
    if "strain" in mode:
        return np.array([[c1, lam, 0], [lam, c1, 0], [0, 0, mu]])
    if "stress" in mode:
        a = E / (1 - nu * nu)
        return a * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    if "3d" in mode:
        D = np.full((3, 3), lam) + np.eye(3) * 2 * mu
        out = np.zeros((6, 6))
        out[:3, :3] = D
        out[3:, 3:] = np.eye(3) * mu
        return out
    raise ValueError("Only for plane-stress, plane-strain, or 3d")


def element_stiffness(domain: VoxelDomain, e_modulus=1.0, poisson_ratio=0.3, plane="strain"):
    """K0 by 2x2(x2) Gauss quadrature at +-h/(2 sqrt 3), weight prod(h/2) (times h_z in 2D)."""
    D = get_D(e_modulus, poisson_ratio, "3d" if domain.dim == 3 else plane)
    siz = domain.element_size
    w = np.prod(siz[: domain.dim] / 2)
    if domain.dim == 2:
        w *= siz[2]
    ndof = domain.elemnodes * domain.dim
    K0 = np.zeros((ndof, ndof))
    for n in domain.node_numbering:
        pos = np.asarray(n, dtype=float) * (siz / 2) / np.sqrt(3)
        B = get_B(domain.eval_shape_fun_der(pos))
        K0 += w * B.T @ D @ B
    return K0


def bc_node_mask(domain: VoxelDomain, ndof: int, bc) -> np.ndarray:
    """Per-node byte mask, bit d set when dof node*ndof+d is Dirichlet."""
    mask = np.zeros(domain.nnodes, dtype=np.uint8)
    if bc is None:
        return mask
    bc = np.unique(np.asarray(bc, dtype=np.int64).ravel())
    if bc.size and (bc.min() < 0 or bc.max() >= domain.nnodes * ndof):
        raise IndexError("boundary condition dof index out of range")
    np.bitwise_or.at(mask, bc // ndof, (1 << (bc % ndof)).astype(np.uint8))
    return mask


class StiffnessMatrix:
    """Matrix-free K(x) on the device with pyMOTO's Dirichlet treatment.

    Behaves like the square sparse matrix the reference returns for ``@``/``dot``/``diagonal``/
    ``shape``/``T``; ``tocsr()`` builds the exact CSR arrays (device tensors) and ``to_scipy()`` a
    host ``scipy.sparse.csr_matrix`` (for small problems and tests).
    """

    ndim = 2
    dtype = np.dtype("float64")

    def __init__(self, grid, xe: torch.Tensor):
        self.grid = grid
        self.xe = xe
        self.shape = (grid.n, grid.n)

    # geometry shortcuts
    @property
    def nel3(self):
        return self.grid.nel3

    @property
    def T(self):
        return self

    def conj(self):
        return self

    def _apply(self, u: torch.Tensor) -> torch.Tensor:
        g, lib = self.grid, _lib.lib()
        y = torch.empty_like(u)
        if g.scalar:
            rc = lib.pmt_scalar_apply(*g.nel3, _lib.ptr(g.K0_dev), _lib.ptr(self.xe), _lib.ptr(g.bcmask_dev), g.bcdiag,
                                      _lib.ptr(u), _lib.ptr(y), _lib.stream())
        else:
            rc = lib.pmt_stiffness_apply(*g.nel3, _lib.ptr(g.K0_dev), g.k0_pattern, _lib.ptr(self.xe),
                                         _lib.ptr(g.bcmask_dev), g.bcdiag, _lib.ptr(u), _lib.ptr(y), _lib.stream())
        _lib.check(rc, "stiffness apply")
        return y

    def __matmul__(self, u):
        u = _lib.as_device(u)
        if u.ndim == 1:
            return self._apply(u)
        return torch.stack([self._apply(u[:, k].contiguous()) for k in range(u.shape[1])], dim=1)

    dot = __matmul__

    def apply_reduced(self, u, precision: str = "fp32") -> torch.Tensor:
        """K u in reduced precision, as the mixed-precision V-cycle applies it on the fine grid:
        ``fp32`` (FP32 arithmetic) or ``bf16`` (hex8: BF16 element products on the matrix cores,
        FP32 accumulation).  Returns an FP32 tensor."""
        g, lib = self.grid, _lib.lib()
        if g.scalar:
            raise NotImplementedError("reduced-precision products exist for the elasticity operator only")
        u32 = _lib.as_device(u, dtype=torch.float32)
        x32 = self.xe.to(torch.float32)
        y = torch.empty_like(u32)
        if precision == "fp32":
            rc = lib.pmt_stiffness_apply_f32(*g.nel3, _lib.ptr(g.K0_dev), g.k0_pattern, _lib.ptr(x32),
                                             _lib.ptr(g.bcmask_dev), g.bcdiag, _lib.ptr(u32), _lib.ptr(y),
                                             _lib.stream())
        elif precision == "bf16":
            rc = lib.pmt_stiffness_apply_bf16(*g.nel3, _lib.ptr(g.K0_dev), _lib.ptr(x32), _lib.ptr(g.bcmask_dev),
                                              g.bcdiag, _lib.ptr(u32), _lib.ptr(y), _lib.stream())
        else:
            raise ValueError("precision must be 'fp32' or 'bf16'")
        _lib.check(rc, f"stiffness apply ({precision})")
        return y

    def residual(self, b: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
        """b - K u in one kernel."""
        g, lib = self.grid, _lib.lib()
        r = torch.empty_like(u)
        if g.scalar:
            rc = lib.pmt_scalar_residual(*g.nel3, _lib.ptr(g.K0_dev), _lib.ptr(self.xe), _lib.ptr(g.bcmask_dev),
                                         g.bcdiag, _lib.ptr(b), _lib.ptr(u), _lib.ptr(r), _lib.stream())
        else:
            rc = lib.pmt_stiffness_residual(*g.nel3, _lib.ptr(g.K0_dev), g.k0_pattern, _lib.ptr(self.xe),
                                            _lib.ptr(g.bcmask_dev), g.bcdiag, _lib.ptr(b), _lib.ptr(u),
                                            _lib.ptr(r), _lib.stream())
        _lib.check(rc, "stiffness residual")
        return r

    def diagonal(self) -> torch.Tensor:
        g, lib = self.grid, _lib.lib()
        d = torch.empty(g.n, dtype=torch.float64, device=self.xe.device)
        fn = lib.pmt_scalar_diagonal if g.scalar else lib.pmt_stiffness_diagonal
        _lib.check(fn(*g.nel3, _lib.ptr(g.K0_dev), _lib.ptr(self.xe), _lib.ptr(g.bcmask_dev), g.bcdiag, _lib.ptr(d),
                      _lib.stream()), "stiffness diagonal")
        return d

    def tocsr(self, index_dtype=torch.int32):
        """(indptr, indices, data) on the device, identical to the reference's CSR arrays."""
        g, lib = self.grid, _lib.lib()
        nb = 4 if index_dtype == torch.int32 else 8
        nnz = lib.pmt_csr_nnz(*g.nel3, g.ndof)
        indptr = torch.empty(g.n + 1, dtype=index_dtype, device=self.xe.device)
        indices = torch.empty(nnz, dtype=index_dtype, device=self.xe.device)
        data = torch.empty(nnz, dtype=torch.float64, device=self.xe.device)
        _lib.check(lib.pmt_csr_pattern(*g.nel3, g.ndof, nb, _lib.ptr(indptr), _lib.ptr(indices), _lib.stream()),
                   "csr pattern")
        _lib.check(lib.pmt_csr_values_ndof(*g.nel3, g.ndof, _lib.ptr(g.K0_dev), _lib.ptr(self.xe),
                                           _lib.ptr(g.bcmask_dev), g.bcdiag, nb, _lib.ptr(indptr), _lib.ptr(data),
                                           _lib.stream()), "csr values")
        return indptr, indices, data

    def to_scipy(self):
        import scipy.sparse as sps
        indptr, indices, data = self.tocsr()
        return sps.csr_matrix((data.cpu().numpy(), indices.cpu().numpy(), indptr.cpu().numpy()), shape=self.shape)

    def toarray(self):
        return self.to_scipy().toarray()


class _GridData:
    """Setup-time data shared by every matrix a module produces (grid, K0, Dirichlet mask)."""

    def __init__(self, domain, K0, bc, bcdiag):
        self.domain = domain
        self.dim = domain.dim
        self.nel3 = (domain.nelx, domain.nely, domain.nelz if domain.dim == 3 else 0)
        self.ndof = K0.shape[0] // domain.elemnodes
        self.n = self.ndof * domain.nnodes
        self.K0 = np.ascontiguousarray(K0, dtype=np.float64)
        self.bc = None if bc is None else np.asarray(bc).ravel()
        self.bcdiag = float(bcdiag) if bcdiag is not None else 0.0
        self.bcmask = bc_node_mask(domain, self.ndof, self.bc)
        self.scalar = self.ndof == 1
        if self.scalar:  # one unknown per node: the element matrix as is
            self.K0_ext = self.K0.ravel().copy()
            self.k0_pattern = 0
        else:  # device copy in the extended layout: K0, then its reflection-basis blocks (hex8)
            self.K0_ext = np.zeros(_lib.K0_EXT_LEN, dtype=np.float64)
            self.k0_pattern = int(_lib.load().pmt_k0_extend(self.dim, self.K0.ctypes.data_as(_lib.c_vp),
                                                            self.K0_ext.ctypes.data_as(_lib.c_vp)))
            if self.k0_pattern < 0:
                raise ValueError("element matrix rejected by pmt_k0_extend")
        self._dev = {}

    def _on_device(self, key, arr, dtype):
        dev = torch.cuda.current_device()
        if (key, dev) not in self._dev:
            self._dev[(key, dev)] = _lib.as_device(arr, dtype)
        return self._dev[(key, dev)]

    @property
    def K0_dev(self):
        return self._on_device("K0", self.K0_ext, torch.float64)

    @property
    def bcmask_dev(self):
        return self._on_device("bcmask", self.bcmask, torch.uint8)

    @property
    def free_mask_dev(self):
        return self._on_device("free", (self.bcmask == 0), torch.bool)


class AssembleGeneral(Module):
    """K(x) = sum_e x_e K_e with Dirichlet rows/cols (``assembly.py:18-315``), device-resident."""

    def __init__(self, domain: VoxelDomain, element_matrix, bc=None, bcdiagval=None, matrix_type=None,
                 add_constant=None, reuse_sparsity: bool = True):
        mats = element_matrix if isinstance(element_matrix, (list, tuple)) else [element_matrix]
        if len(mats) != 1:
            raise NotImplementedError("only a single element matrix is supported on the device")
        K0 = np.asarray(mats[0])
        if np.iscomplexobj(K0):
            raise NotImplementedError("complex element matrices are not supported on the device")
        if K0.ndim != 2 or K0.shape[0] != K0.shape[1]:
            raise NotImplementedError("only square element matrices are supported on the device")
        if K0.shape[0] % domain.elemnodes != 0:
            raise ValueError("Number of rows in element matrix should be a multiple of the number of nodes per element")
        if K0.shape[0] // domain.elemnodes not in (1, domain.dim):
            raise NotImplementedError("the device operators take dim unknowns per node (elasticity) or one "
                                      "(scalar problems)")
        if domain.dim not in (2, 3):
            raise NotImplementedError("1D domains are not supported")
        if add_constant is not None:
            raise NotImplementedError("add_constant is not supported by the matrix-free operator")
        if bc is not None and bcdiagval is None:
            bcdiagval = np.max(K0)
        self.elmat = [K0]
        self.nel = domain.nel
        self.grid = _GridData(domain, K0, bc, bcdiagval)
        self.bc = self.grid.bc
        self.bcdiagval = bcdiagval
        self.matrix_type = matrix_type
        self.reuse_sparsity = reuse_sparsity

    def __call__(self, x):
        xe = _lib.as_device(x)
        if xe.numel() != self.nel:
            raise ValueError(f"Input vector wrong size ({xe.numel()}), must be equal to #nel ({self.nel})")
        return StiffnessMatrix(self.grid, xe.reshape(-1))

    def _sensitivity(self, dgdmat):
        if dgdmat is None:
            return [None]
        if not isinstance(dgdmat, DyadicMatrix):
            raise NotImplementedError("the device assembly back-propagates DyadicMatrix sensitivities only")
        g, lib = self.grid, _lib.lib()
        xs = self.sig_in[0].state
        dx = torch.zeros(self.nel, dtype=torch.float64, device=_lib.device())
        for k, (u, v, fac) in enumerate(dgdmat.terms()):
            if g.scalar:
                rc = lib.pmt_scalar_sens_contract(*g.nel3, _lib.ptr(g.K0_dev), _lib.ptr(g.bcmask_dev), _lib.ptr(u),
                                                  _lib.ptr(v), float(fac), _lib.ptr(dx), int(k > 0), _lib.stream())
            else:
                rc = lib.pmt_sens_contract_k0(*g.nel3, _lib.ptr(g.K0_dev), g.k0_pattern, _lib.ptr(g.bcmask_dev),
                                              _lib.ptr(u), _lib.ptr(v), float(fac), _lib.ptr(dx), int(k > 0),
                                              _lib.stream())
            _lib.check(rc, "sensitivity contraction")
        if isinstance(xs, np.ndarray):
            return dx.cpu().numpy().reshape(xs.shape)
        return dx.reshape(xs.shape)


class AssembleStiffness(AssembleGeneral):
    """Linear-elastic stiffness K(x) = sum_e x_e K0 for Q4/hex8 (``assembly.py:407-463``)."""

    def __init__(self, domain: VoxelDomain, *args, e_modulus: float = 1.0, poisson_ratio: float = 0.3,
                 plane="strain", **kwargs):
        self.E, self.nu = e_modulus, poisson_ratio
        self.stiffness_element = element_stiffness(domain, e_modulus, poisson_ratio, plane)
        super().__init__(domain, self.stiffness_element, *args, **kwargs)


def poisson_element(domain: VoxelDomain, material_property=1.0):
    """Element matrix of the Poisson equation, P0 = sum_gp w k grad(N)^T grad(N) with 2x2(x2) Gauss points
    at +-h/(2 sqrt 3) and weight prod(h/2); in 2D k is multiplied by the thickness h_z
    (``assembly.py:535-560``)."""
    siz = domain.element_size
    w = np.prod(siz[: domain.dim] / 2)
    k = np.prod(np.atleast_1d(material_property)) * (siz[2] if domain.dim == 2 else 1.0)
    P0 = np.zeros((domain.elemnodes, domain.elemnodes))
    for n in domain.node_numbering:
        pos = np.asarray(n, dtype=float) * (siz / 2) / np.sqrt(3)
        dN = domain.eval_shape_fun_der(pos)
        P0 += w * k * dN.T @ dN
    return P0


class AssemblePoisson(AssembleGeneral):
    """Scalar conduction / permittivity matrix P(x) = sum_e x_e P0 for Q4/hex8, one unknown per node
    (``assembly.py:523-560``); applied, solved and differentiated by the scalar kernels."""

    def __init__(self, domain: VoxelDomain, *args, material_property: float = 1.0, **kwargs):
        self.material_property = material_property
        self.poisson_element = poisson_element(domain, material_property)
        super().__init__(domain, self.poisson_element, *args, **kwargs)
