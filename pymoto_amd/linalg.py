"""``LinSolve`` on the device (``pymoto/modules/linalg.py:121-219``).

Forward: wrap the solver in ``LDAWrapper`` (tolerance 5x the CG tolerance, ``linalg.py:185-189``),
``update(K)`` and ``u = solve(f, x0=previous u)`` (warm start, ``linalg.py:192-195``).  Adjoint:
``lam = solve(dfdv, trans='T')`` - for compliance the LDAS projection reproduces lam from the stored
solution without a second solve - and the matrix sensitivity ``DyadicMatrix(-lam, u)``
(``linalg.py:199-219``).  The matrix must be the ``StiffnessMatrix`` of ``AssembleStiffness``.
"""
import numpy as np

from . import _lib
from .assembly import StiffnessMatrix
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
