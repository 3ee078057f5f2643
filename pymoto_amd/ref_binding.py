"""Reference-side binding: a pyMOTO ``LinearSolver`` that solves on the MI355X through the C-ABI only.

This is the stub a pyMOTO maintainer would add as ``pymoto/solvers/mi355x.py`` (it is reproduced in the
integration notes, there deriving from ``pymoto.solvers.LinearSolver``).  It needs nothing but
``ctypes``, numpy, the HIP runtime (``libamdhip64.so``) and ``libpymoto_amd.so`` - no torch - and
keeps pyMOTO's solver contract (``pymoto/solvers/solvers.py:6-50``):

* ``update(A)`` is called by ``LinSolve`` with the assembled matrix of every design iteration
  (``pymoto/modules/linalg.py:191-192``).  A matrix-free operator needs the element factors that
  ``A`` was assembled from, so the solver is given the ``Signal`` that feeds ``AssembleStiffness``
  and uploads its current state at each ``update``; ``A`` itself is only checked for its shape.
* ``solve(rhs, x0=None, trans='N')`` (``solvers.py:36-50``) honours ``x0`` (``x0_is_zero = 0``) and
  ``trans`` ('N', 'T' and 'H' are the same solve for a real symmetric stiffness matrix); a 2D
  right-hand side is solved column by column.
* Device buffers (densities, Dirichlet mask, one right-hand side and one solution) and one HIP stream
  are allocated once and freed by ``close()`` (also called on garbage collection), with the multigrid
  handle.  The HIP runtime calls go through ``libpymoto_amd.so``'s own handle, so they reach the
  runtime the library is linked against.
"""
import ctypes
import ctypes.util
import os

import numpy as np

_HIP_H2D, _HIP_D2H = 1, 2
_LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libpymoto_amd.so")
_vp, _i, _d, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t


def _load():
    pmt = ctypes.CDLL(_LIBPATH)
    # the HIP entry points are looked up through the library's own handle, so they come from the HIP
    # runtime libpymoto_amd.so is linked against (one runtime per process)
    hip = pmt
    hip.hipMalloc.argtypes = [ctypes.POINTER(_vp), _sz]
    hip.hipFree.argtypes = [_vp]
    hip.hipMemcpy.argtypes = [_vp, _vp, _sz, _i]
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(_vp)]
    hip.hipStreamSynchronize.argtypes = [_vp]
    hip.hipStreamDestroy.argtypes = [_vp]
    pmt.pmt_last_error.restype = ctypes.c_char_p
    pmt.pmt_mg_create.argtypes = [_i] * 4 + [_vp] * 4
    pmt.pmt_mg_destroy.argtypes = [_vp]
    pmt.pmt_mg_update.argtypes = [_vp, _vp, _vp, _d, _vp]
    pmt.pmt_pcg_solve.argtypes = [_vp, _vp, _vp, _i, _d, _i, _i, _vp, _vp, _vp, _vp]
    return hip, pmt


class SolverMGPCG:
    """CG preconditioned by geometric multigrid on the GPU, for ``AssembleStiffness`` systems on a
    ``VoxelDomain`` (same damped-Jacobi V(steps, steps) cycle and Galerkin coarse operators as
    ``GeometricMultigrid`` + ``CG``, ``pymoto/solvers/iterative.py:124-403``).

    Args:
        domain: the ``VoxelDomain`` (``nelx, nely, nelz, dim, nnodes``).
        element_matrix: ``AssembleStiffness.stiffness_element`` (the K0 the matrix is assembled from).
        bc: Dirichlet dof indices, as given to ``AssembleStiffness``.
        bcdiagval: the Dirichlet diagonal, ``AssembleGeneral.bcdiagval`` (max of K0 by default).
        design: the ``Signal`` whose state (element factors x_e, e.g. the SIMP output) feeds the assembly.
        levels, omega, steps: grids of the hierarchy (fine included), Jacobi damping, sweeps per level.
        tol, maxit, restart: ``CG`` stopping rule (``iterative.py:306-313``).
    """

    def __init__(self, domain, element_matrix, bc, bcdiagval, design, levels=4, omega=0.6, steps=2, tol=1e-8,
                 maxit=10000, restart=50):
        self._hip, self._pmt = _load()
        self.domain, self.design = domain, design
        self.tol, self.maxit, self.restart = float(tol), int(maxit), int(restart)
        self.iterations, self.relres = 0, np.nan
        self.h = _vp()
        self._bufs = []
        dim = domain.dim
        self.n = domain.nnodes * dim
        om = np.full(levels - 1, omega, dtype=np.float64)
        st = np.full(levels - 1, steps, dtype=np.int32)
        K0 = np.ascontiguousarray(element_matrix, dtype=np.float64)
        self._check(self._pmt.pmt_mg_create(domain.nelx, domain.nely, domain.nelz if dim == 3 else 0, levels,
                                            om.ctypes.data, st.ctypes.data, K0.ctypes.data, ctypes.byref(self.h)))
        mask = np.zeros(domain.nnodes, dtype=np.uint8)
        bc = np.asarray(bc, dtype=np.int64).ravel()
        np.bitwise_or.at(mask, bc // dim, (1 << (bc % dim)).astype(np.uint8))
        self.bcdiag = float(bcdiagval)
        self.d_mask = self._alloc(mask.nbytes)
        self._h2d(self.d_mask, mask)
        self.d_x = self._alloc(domain.nel * 8)
        self.d_b = self._alloc(self.n * 8)
        self.d_u = self._alloc(self.n * 8)
        # the solver's own (blocking) stream: the legacy default stream's synchronous copies order
        # themselves with it, and the copies back wait for it explicitly
        self.stream = _vp()
        if self._hip.hipStreamCreate(ctypes.byref(self.stream)) != 0:
            raise RuntimeError("hipStreamCreate failed")

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"pymoto_amd status {rc}: {self._pmt.pmt_last_error().decode(errors='replace')}")

    def _alloc(self, nbytes):
        p = _vp()
        if self._hip.hipMalloc(ctypes.byref(p), _sz(nbytes)) != 0:
            raise MemoryError(f"hipMalloc of {nbytes} bytes failed")
        self._bufs.append(p)
        return p

    def _h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        if self._hip.hipMemcpy(dst, arr.ctypes.data, _sz(arr.nbytes), _HIP_H2D) != 0:
            raise RuntimeError("hipMemcpy host to device failed")

    def _d2h(self, arr, src):
        if self._hip.hipStreamSynchronize(self.stream) != 0:
            raise RuntimeError("hipStreamSynchronize failed")
        if self._hip.hipMemcpy(arr.ctypes.data, src, _sz(arr.nbytes), _HIP_D2H) != 0:
            raise RuntimeError("hipMemcpy device to host failed")

    def update(self, A):
        if A is not None and tuple(A.shape) != (self.n, self.n):
            raise ValueError(f"matrix of shape {A.shape} does not belong to this domain ({self.n} dofs)")
        x = np.ascontiguousarray(np.asarray(self.design.state, dtype=np.float64).ravel())
        if x.size != self.domain.nel:
            raise ValueError("the design signal must hold one factor per element")
        self._h2d(self.d_x, x)  # the current design, every update
        self._check(self._pmt.pmt_mg_update(self.h, self.d_x, self.d_mask, self.bcdiag, self.stream))

    def _solve1(self, b, x0):
        self._h2d(self.d_b, np.asarray(b, dtype=np.float64))
        if x0 is not None:
            self._h2d(self.d_u, np.asarray(x0, dtype=np.float64))
        it, rr = ctypes.c_int(), ctypes.c_double()
        self._check(self._pmt.pmt_pcg_solve(self.h, self.d_b, self.d_u, int(x0 is None), self.tol, self.maxit,
                                            self.restart, ctypes.byref(it), ctypes.byref(rr), None, self.stream))
        self.iterations, self.relres = it.value, rr.value
        out = np.empty(self.n)
        self._d2h(out, self.d_u)
        return out

    def solve(self, rhs, x0=None, trans="N"):
        if trans not in ("N", "T", "H"):
            raise TypeError("Only N, T, or H transposition is possible")
        rhs = np.asarray(rhs)
        if np.iscomplexobj(rhs):
            raise TypeError("real-valued solver")
        if rhs.ndim == 1:
            return self._solve1(rhs, x0)
        x0c = (lambda k: None) if x0 is None else (lambda k: np.asarray(x0)[:, k])
        return np.stack([self._solve1(rhs[:, k], x0c(k)) for k in range(rhs.shape[1])], axis=1)

    def close(self):
        if getattr(self, "_hip", None) is None:
            return
        self._hip.hipDeviceSynchronize()
        if self.h:
            self._pmt.pmt_mg_destroy(self.h)
            self.h = _vp()
        if getattr(self, "stream", None):
            self._hip.hipStreamDestroy(self.stream)
            self.stream = _vp()
        for p in self._bufs:
            self._hip.hipFree(p)
        self._bufs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
