"""Solver base class and the linear-dependency-aware wrapper on the device.

Mirrors ``pymoto/solvers/solvers.py``: ``LinearSolver`` (``update``/``solve``/``residual``,
lines 6-85) and ``LDAWrapper`` (lines 99-306).  LDAS keeps normalised pairs (x_k, b_k = A x_k) of
earlier solutions; a new right-hand side is first projected onto them (modified Gram-Schmidt),
the residual of that reconstruction is checked with one operator application, and the inner
solver only runs when it exceeds ``tol`` - which is why the compliance adjoint costs no solve.
Dirichlet dofs (rows with only a diagonal entry) are solved by division (lines 174-176).  All
vector work runs in the HIP library; the scalars of the projections come back to the host.
"""
import warnings

import numpy as np
import torch

from .. import _lib
from ..assembly import StiffnessMatrix


class LinearSolver:
    defined = True

    def __init__(self, A=None):
        if A is not None:
            self.update(A)

    def update(self, A):
        raise NotImplementedError("Solver not implemented")

    def solve(self, rhs, x0=None, trans="N"):
        raise NotImplementedError("Solver not implemented")

    @staticmethod
    def residual(A, x, b, trans="N"):
        """||A x - b|| / ||b|| per column (1 for a zero b), ``solvers.py:52-85``."""
        x, b = _lib.as_device(x), _lib.as_device(b)
        cols = [(x, b)] if x.ndim == 1 else [(x[:, k].contiguous(), b[:, k].contiguous()) for k in range(x.shape[1])]
        out = []
        for xc, bc in cols:
            r = A.residual(bc, xc)
            bn = np.sqrt(_dot(bc, bc))
            out.append(np.sqrt(_dot(r, r)) / (bn if bn > 0 else 1.0))
        return out[0] if x.ndim == 1 else np.array(out)


def _dot(a, b) -> float:
    lib = _lib.lib()
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    _lib.check(lib.pmt_dot(a.numel(), _lib.ptr(a), _lib.ptr(b), _lib.ptr(out), _lib.ptr(_lib.reduce_work()),
                           _lib.stream()), "dot")
    return float(out.item())


def _dots(pairs):
    """The dot products of several vector pairs with one host synchronisation (numpy array)."""
    lib = _lib.lib()
    out = torch.empty(len(pairs), dtype=torch.float64, device=pairs[0][0].device)
    for k, (a, b) in enumerate(pairs):
        _lib.check(lib.pmt_dot(a.numel(), _lib.ptr(a), _lib.ptr(b), _lib.ptr(out[k:k + 1]),
                               _lib.ptr(_lib.reduce_work()), _lib.stream()), "dot")
    return out.cpu().numpy()


def _axpy(a, x, y):
    """y += a x (in place)"""
    _lib.check(_lib.lib().pmt_axpby(y.numel(), float(a), _lib.ptr(x), 1.0, _lib.ptr(y), _lib.stream()), "axpby")


def _scale(a, y):
    _lib.check(_lib.lib().pmt_axpby(y.numel(), 0.0, None, float(a), _lib.ptr(y), _lib.stream()), "scale")


class LDAWrapper(LinearSolver):
    """Linear dependency aware solver (Koppen et al. 2022), ``solvers.py:99-306``, real symmetric case."""

    def __init__(self, solver: LinearSolver, tol=1e-7, A=None, symmetric=None, hermitian=None):
        self.solver = solver
        self.tol = tol
        self.x_stored, self.b_stored = [], []
        self.A = None
        self.symmetric = symmetric
        self.hermitian = hermitian
        self._did_solve = False
        self._last_rtol = 0.0
        super().__init__(A)

    def update(self, A, skip_inner_update: bool = False):
        self.A = A
        self.x_stored.clear()
        self.b_stored.clear()
        if not skip_inner_update:
            self.solver.update(A)

    def _split(self, v):
        """(v with Dirichlet entries zeroed, v[bc]/bcdiag on Dirichlet entries and 0 elsewhere)."""
        g = self.A.grid
        rl, sol = torch.empty_like(v), torch.empty_like(v)
        _lib.check(_lib.lib().pmt_bc_split_ndof(*g.nel3, g.ndof, _lib.ptr(g.bcmask_dev), g.bcdiag if g.bcdiag else 1.0,
                                                _lib.ptr(v), _lib.ptr(rl), _lib.ptr(sol), _lib.stream()), "bc split")
        return rl, sol

    def _solve_cols(self, cols, x0s):
        """LDAS for a list of right-hand-side columns solved together (``solvers.py:160-248``):
        reconstruct every column from the database, check each residual, solve the columns above
        ``tol`` in one call of the inner solver (a block solve for several), add them to the database."""
        A, g = self.A, self.A.grid
        sols, rls, rtols = [], [], []
        for rhs in cols:
            rl, sol = self._split(rhs)
            for x, b in zip(self.x_stored, self.b_stored):
                alpha = _dot(rl, b) / _dot(b, b)
                _axpy(-alpha, b, rl)
                _axpy(alpha, x, sol)
            if self.x_stored or not isinstance(A, StiffnessMatrix):
                r = A.residual(rhs, sol)
            else:
                # empty database: sol holds only the Dirichlet values rhs[bc] / bcdiag, the operator's
                # Dirichlet rows are bcdiag times the identity and its Dirichlet columns vanish, so
                # rhs - A sol is rl (up to the rounding of rhs[bc] / bcdiag * bcdiag): no operator pass
                r = rl
            bb, rr = _dots([(rhs, rhs), (r, r)])
            bn = np.sqrt(bb)
            rtols.append(np.sqrt(rr) / (bn if bn > 0 else 1.0))
            sols.append(sol)
            rls.append(rl)
        did = [rt > self.tol for rt in rtols]
        todo = [k for k in range(len(cols)) if did[k]]
        if todo:
            x0l = None
            if x0s is not None:
                x0l = []
                for k in todo:
                    v, _ = self._split(_lib.as_device(x0s[k]))  # new tensors: x0 is not modified
                    for x in self.x_stored:
                        _axpy(-_dot(v, x) / _dot(x, x), x, v)
                    x0l.append(v)
            solve1 = getattr(self.solver, "_solve_1rhs", None)  # the device CG
            if len(todo) == 1 and x0l is not None and solve1 is not None:
                # x0l[0] is this wrapper's own projected copy: the CG may iterate in it
                xnews = [solve1(rls[todo[0]], x0l[0], x0_scratch=True)]
            elif len(todo) == 1:
                xnews = [self.solver.solve(rls[todo[0]], x0=None if x0l is None else x0l[0])]
            else:
                X = self.solver.solve(torch.stack([rls[k] for k in todo], dim=1),
                                      x0=None if x0l is None else torch.stack(x0l, dim=1))
                xnews = [X[:, i].contiguous() for i in range(len(todo))]
            for k, xnew in zip(todo, xnews):
                _lib.check(_lib.lib().pmt_axpy_free_ndof(*g.nel3, g.ndof, _lib.ptr(g.bcmask_dev), 1.0,
                                                         _lib.ptr(xnew), _lib.ptr(sols[k]), _lib.stream()),
                           "axpy_free")
            for xnew in xnews:
                xadd, _ = self._split(xnew)
                badd, _ = self._split(A @ xnew)
                for x, b in zip(self.x_stored, self.b_stored):
                    beta = _dot(badd, b) / _dot(b, b)
                    _axpy(-beta, b, badd)
                    _axpy(-beta, x, xadd)
                bnrm = np.sqrt(_dot(badd, badd))
                if np.isfinite(bnrm) and bnrm > 0:
                    _scale(1.0 / bnrm, badd)
                    _scale(1.0 / bnrm, xadd)
                    self.x_stored.append(xadd)
                    self.b_stored.append(badd)
        return sols, rtols, did

    def solve(self, rhs, x0=None, trans="N"):
        if trans.upper() not in ("N", "T", "H"):
            raise TypeError("Only N, T, or H transposition is possible")
        rhs = _lib.as_device(rhs)
        if rhs.ndim == 1:
            sols, rt, did = self._solve_cols([rhs], None if x0 is None else [x0])
            self._last_rtol, self._did_solve = rt[0], did[0]
            return sols[0]
        cols = [rhs[:, k].contiguous() for k in range(rhs.shape[1])]
        x0s = None if x0 is None else [_lib.as_device(x0)[:, k].contiguous() for k in range(rhs.shape[1])]
        sols, rt, did = self._solve_cols(cols, x0s)
        self._last_rtol, self._did_solve = np.array(rt), np.array(did)
        return torch.stack(sols, dim=1)
