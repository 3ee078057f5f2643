"""``LinSolve`` on the device (``pymoto/modules/linalg.py:121-219``).

Forward: wrap the solver in ``LDAWrapper`` (tolerance 5x the CG tolerance, ``linalg.py:185-189``),
``update(K)`` and ``u = solve(f, x0=previous u)`` (warm start, ``linalg.py:192-195``).  Adjoint:
``lam = solve(dfdv, trans='T')`` - for compliance the LDAS projection reproduces lam from the stored
solution without a second solve - and the matrix sensitivity ``DyadicMatrix(-lam, u)``
(``linalg.py:199-219``).  The matrix must be the ``StiffnessMatrix`` of ``AssembleStiffness``.
"""
import numpy as np
import torch

from . import _lib
from .assembly import StiffnessMatrix, _GridData
from .core import Module
from .dyadic import DyadicMatrix
from .solvers import LDAWrapper, auto_determine_solver


class LinSolve(Module):
    use_lda_solver = True

    def __init__(self, dep_tol: float = 1e-5, hermitian: bool = None, symmetric: bool = None,
                 positive_definite: bool = None, solver=None):
        self.dep_tol = dep_tol
        self.ishermitian = hermitian
        self.issymmetric = symmetric
        self.ispositivedefinite = positive_definite
        self.solver = solver
        self.u = None

    def __call__(self, mat, rhs):
        if not isinstance(mat, StiffnessMatrix):
            raise TypeError("the device LinSolve solves StiffnessMatrix systems produced by AssembleStiffness")
        if np.iscomplexobj(rhs) if isinstance(rhs, np.ndarray) else False:
            raise TypeError("Complex right-hand-side for a real-valued sparse matrix is not supported.")
        self.ishermitian = True if self.ishermitian is None else self.ishermitian
        if self.solver is None:
            self.solver = auto_determine_solver(mat)
        if not isinstance(self.solver, LDAWrapper) and self.use_lda_solver:
            kw = dict(hermitian=True, symmetric=True)
            if hasattr(self.solver, "tol"):
                kw["tol"] = self.solver.tol * 5
            self.solver = LDAWrapper(self.solver, **kw)
        self.solver.update(mat)
        self.u = self.solver.solve(_lib.as_device(rhs), x0=self.u)
        return self.u

    def _sensitivity(self, dfdv):
        mat, rhs = self.get_input_states()
        lam = self.solver.solve(_lib.as_device(dfdv), trans="T")
        if self.u.ndim > 1:
            dmat = DyadicMatrix([lam[:, k].contiguous() for k in range(lam.shape[1])],
                                [self.u[:, k].contiguous() for k in range(self.u.shape[1])], fac=-1.0)
        else:
            dmat = DyadicMatrix(lam, self.u, fac=-1.0)
        db = lam.cpu().numpy() if isinstance(rhs, np.ndarray) else lam
        return dmat, db


class SystemOfEquations(Module):
    """Solve with prescribed unknowns (``pymoto/modules/linalg.py:222-349``): the rows ``free`` carry the
    applied right-hand side b_f, the rows ``prescribed`` the given values x_p, and

        A_ff x_f = b_f - A_fp x_p,   b_p = A_pf x_f + A_pp x_p   (outputs: the full x and b).

    On the device nothing is partitioned: A_ff is the same matrix-free operator with the prescribed
    unknowns added to its Dirichlet set (their rows and columns drop out exactly as in the partition),
    A_fp x_p and b_p come from one operator application each on full-length vectors, and the solve is
    an inner ``LinSolve`` (``**kwargs`` are its arguments; LDAS warm start across calls).
    Adjoint (``linalg.py:312-349``): lam_f = -A_ff^-T (dg/dx_f + A_fp dg/db_p), lam_p = dg/db_p, dg/dA =
    lam (x) x, dg/db_f = dg/db_f - lam_f, dg/dx_p = (A lam)_p + dg/dx_p.
    """

    def __init__(self, free=None, prescribed=None, **kwargs):
        if free is None and prescribed is None:
            raise ValueError("Either prescribed or free indices must be provided")
        self.linsolve = LinSolve(**kwargs)
        self.f = None if free is None else np.asarray(free, dtype=np.int64).ravel()
        self.p = None if prescribed is None else np.asarray(prescribed, dtype=np.int64).ravel()
        self._free_grid = None
        self._key = None

    def _indices(self, n):
        if self.f is None:
            self.f = np.setdiff1d(np.arange(n), self.p)
        if self.p is None:
            self.p = np.setdiff1d(np.arange(n), self.f)
        if self.f.size + self.p.size != n:
            raise ValueError("Size of free and prescribed indices must match the matrix size")
        dev = _lib.device()
        return torch.as_tensor(self.f, device=dev), torch.as_tensor(self.p, device=dev)

    def _free_operator(self, A):
        """A with the prescribed unknowns as extra Dirichlet rows/cols: A_ff on the free rows."""
        g = A.grid
        key = (id(g), self.p.tobytes())
        if self._key != key:
            bc = self.p if g.bc is None else np.union1d(g.bc, self.p)
            bcdiag = g.bcdiag if g.bc is not None and g.bcdiag else float(np.max(g.K0))
            self._free_grid, self._key = _GridData(g.domain, g.K0, bc, bcdiag), key
        return StiffnessMatrix(self._free_grid, A.xe)

    @staticmethod
    def _sub(a, b):
        """a - b (same shape, contiguous) in the HIP library."""
        out = b.contiguous().clone()
        _lib.check(_lib.lib().pmt_axpby(out.numel(), 1.0, _lib.ptr(a.contiguous()), -1.0, _lib.ptr(out),
                                        _lib.stream()), "axpby")
        return out

    @staticmethod
    def _add(a, b):
        out = b.contiguous().clone()
        _lib.check(_lib.lib().pmt_axpby(out.numel(), 1.0, _lib.ptr(a.contiguous()), 1.0, _lib.ptr(out),
                                        _lib.stream()), "axpby")
        return out

    def __call__(self, A, bf, xp):
        if not isinstance(A, StiffnessMatrix):
            raise TypeError("the device SystemOfEquations takes the StiffnessMatrix of an assembly module")
        n = A.shape[0]
        bf, xp = _lib.as_device(bf), _lib.as_device(xp)
        assert bf.shape[0] + xp.shape[0] == n, "Dimensions of applied force and displacement must match matrix"
        assert bf.ndim == xp.ndim, "Number of loadcases for applied force and displacement must match"
        fi, pi = self._indices(n)
        self._fi, self._pi, self.A = fi, pi, A
        xbar = torch.zeros((n,) + tuple(bf.shape[1:]), dtype=torch.float64, device=bf.device)
        xbar.index_copy_(0, pi, xp)
        rhs = torch.zeros_like(xbar)
        rhs.index_copy_(0, fi, self._sub(bf, (A @ xbar).index_select(0, fi)))  # b_f - A_fp x_p
        Aff = self._free_operator(A)
        xf = self.linsolve(Aff, rhs)  # zero on the prescribed rows
        self.x = xf.clone()
        self.x.index_copy_(0, pi, xp)
        b = (A @ self.x).contiguous()
        b.index_copy_(0, fi, bf)
        return self.x, b

    def _sensitivity(self, dgdx, dgdb):
        A, fi, pi = self.A, self._fi, self._pi
        dgdx = None if dgdx is None else _lib.as_device(dgdx)
        dgdb = None if dgdb is None else _lib.as_device(dgdb)
        load = torch.zeros_like(self.x)
        if dgdx is not None:
            load.index_copy_(0, fi, dgdx.index_select(0, fi))
        zbar = None
        if dgdb is not None:
            zbar = torch.zeros_like(self.x)
            zbar.index_copy_(0, pi, dgdb.index_select(0, pi))
            cpl = (A @ zbar).index_select(0, fi)  # A_fp dg/db_p
            load.index_copy_(0, fi, self._add(cpl, load.index_select(0, fi)))
        lam = self.linsolve.solver.solve(load, trans="T")  # +A_ff^-1 load, zero on the prescribed rows
        _lib.check(_lib.lib().pmt_axpby(lam.numel(), 0.0, None, -1.0, _lib.ptr(lam), _lib.stream()), "scale")
        if zbar is not None:
            lam = self._add(zbar, lam)  # lam_p = dg/db_p
        if self.x.ndim > 1:
            dgdA = DyadicMatrix([lam[:, k].contiguous() for k in range(lam.shape[1])],
                                [self.x[:, k].contiguous() for k in range(self.x.shape[1])])
        else:
            dgdA = DyadicMatrix(lam, self.x)
        lam_f = lam.index_select(0, fi)
        dgdbf = self._sub(dgdb.index_select(0, fi), lam_f) if dgdb is not None else self._sub(
            torch.zeros_like(lam_f), lam_f)
        dgdup = (A @ lam).index_select(0, pi)  # A_pf lam_f + A_pp dg/db_p
        if dgdx is not None:
            dgdup = self._add(dgdx.index_select(0, pi), dgdup)
        return dgdA, dgdbf, dgdup
