"""Oracle: one compliance design iteration, filter -> SIMP -> assembly -> solve -> c -> dc/dx.

Test infrastructure only.  Restates the module chain of
``examples/topology_optimization/ex_compliance.py:84-110`` /
``ex_compliance_multigrid.py:84-131`` with ``LinSolve`` + ``LDAWrapper`` semantics
(``linalg.py:164-219``, ``solvers.py:140-306``) and the compliance adjoint
(``generic.py:196-226``, ``assembly.py:298-315``, ``generic.py:101-140``, ``filter.py:214-220``).

Pinned end-to-end against the reference's own outputs in ``tests/golden/*.npz``.
"""
import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from . import fe
from .filter import ConvFilter, filter_weights
from .solvers import CG, LDAS, DirectLU, build_mg


def design_iteration(nel, x, bc, f, radius, xmin=1e-9, penal=3.0, plane="strain", solver="direct",
                     mg_levels=4, omega=0.6, smooth_steps=2, tol=1e-8, filter_bcs=None, E=1.0, nu=0.3,
                     x0=None, return_all=False):
    nelx, nely, nelz = nel
    dim = fe.grid_dim(nelx, nely, nelz)
    w = filter_weights(radius, nel)
    filt = ConvFilter(nel, w, filter_bcs or (("symmetric",) * 2,) * 3)
    xf = filt.forward(x)
    xs = fe.simp(xf, xmin, penal)
    K0 = fe.element_stiffness(dim, E=E, nu=nu, plane=plane)
    asm = fe.Assembly(nelx, nely, nelz, K0, bc=bc)
    K = asm.matrix(xs)
    if solver == "direct":
        inner = DirectLU()
        lda = LDAS(inner, tol=1e-7)
    else:
        mg = build_mg(nel, mg_levels, omega=omega, smooth_steps=smooth_steps)
        inner = CG(preconditioner=mg, tol=tol)
        lda = LDAS(inner, tol=5 * tol)
    lda.update(K)
    u = lda.solve(f, x0=x0)
    c = float(f @ u)
    lam = lda.solve(1.0 * f)  # LinSolve._sensitivity with dc = 1 (EinSum adjoint: dc * f)
    dxs = asm.sens_contract([-lam], [u])
    dxf = dxs * fe.dsimp(xf, xmin, penal)
    dx = filt.adjoint(dxf)
    out = dict(xf=xf, xsimp=xs, u=u, c=c, dcdxsimp=dxs, dcdxf=dxf, dcdx=dx)
    if solver != "direct":
        out["iterations"] = inner.iterations
        out["history"] = np.array(inner.history)
        out["adjoint_solved"] = bool(np.any(lda.did_solve))
    if return_all:
        out.update(K=K, K0=K0, asm=asm, filt=filt)
    return out


def heaviside(x, beta, eta):
    """Heaviside projection of ``examples/topology_optimization/ex_compliance_robust.py:139`` (the
    MathExpression ``(tanh(b*eta) + tanh(b*(x-eta))) / (tanh(b*eta) + tanh(b*(1-eta)))``) and its
    derivative d/dx = b (1 - tanh^2(b (x - eta))) / (tanh(b*eta) + tanh(b*(1-eta)))."""
    den = np.tanh(beta * eta) + np.tanh(beta * (1.0 - eta))
    t = np.tanh(beta * (x - eta))
    return (np.tanh(beta * eta) + t) / den, beta * (1.0 - t * t) / den


def robust_iteration(nel, x, bc, F, radius, etas=(0.3, 0.5, 0.7), beta=8.0, xmin=1e-9, penal=3.0, E=1.0, nu=0.3):
    """Robust multi-load design iteration: filter -> Heaviside at each eta -> SIMP -> K -> solve the k load
    cases F (ndof, k) with the direct solve -> c = sum_j f_j.u_j (EinSum 'ij,ij->j' then 'i->') -> dc/dx
    through the adjoint chain (self-adjoint: lambda = u, ``linalg.py:199-219``).  One entry per eta:
    xp (projected), U, c and dcdx.  Chain of ``ex_compliance_robust.py:132-149`` with the multi-load
    compliance of ``ex_compliant_mechanism.py:83``."""
    nelx, nely, nelz = nel
    dim = fe.grid_dim(nelx, nely, nelz)
    filt = ConvFilter(nel, filter_weights(radius, nel), (("symmetric",) * 2,) * 3)
    xf = filt.forward(x)
    K0 = fe.element_stiffness(dim, E=E, nu=nu)
    asm = fe.Assembly(nelx, nely, nelz, K0, bc=bc)
    out = []
    for eta in etas:
        xp, dxp = heaviside(xf, beta, eta)
        xs = fe.simp(xp, xmin, penal)
        lu = DirectLU()
        lu.update(asm.matrix(xs))
        U = lu.solve(F)
        c = float(np.sum(F * U))
        dxs = asm.sens_contract([-U[:, j] for j in range(F.shape[1])], [U[:, j] for j in range(F.shape[1])])
        dx = filt.adjoint(dxs * fe.dsimp(xp, xmin, penal) * dxp)
        out.append(dict(eta=eta, xf=xf, xp=xp, xsimp=xs, U=U, c=c, dcdx=dx))
    return out


class SystemOfEquations:
    """Partitioned solve with prescribed unknowns (``SystemOfEquations``, pymoto/modules/linalg.py:222-349),
    dense-free numpy/scipy restatement with the direct solve: A_ff x_f = b_f - A_fp x_p, b_p = A_pf x_f +
    A_pp x_p; outputs the full x and b.  ``sensitivity(dgdx, dgdb)`` returns (lam, x) of the matrix dyad
    lam (x) x, dg/db_f and dg/dx_p (``linalg.py:312-349``).  Pinned by finite differences of a scalar
    response through all three inputs and by the identities A x = b, x_p = prescribed, b_f = applied."""

    def __init__(self, free=None, prescribed=None):
        if free is None and prescribed is None:
            raise ValueError("Either prescribed or free indices must be provided")
        self.f, self.p = free, prescribed

    def forward(self, A, bf, xp):
        A = sps.csr_matrix(A)
        n = A.shape[0]
        if self.f is None:
            self.f = np.setdiff1d(np.arange(n), self.p)
        if self.p is None:
            self.p = np.setdiff1d(np.arange(n), self.f)
        f, p = self.f, self.p
        self.Aff = A[f][:, f].tocsc()
        self.Afp = A[f][:, p]
        self.App = A[p][:, p]
        self.lu = spla.splu(self.Aff)
        x = np.zeros((n,) + bf.shape[1:])
        x[p] = xp
        x[f] = self.lu.solve(bf - self.Afp @ xp)
        b = np.zeros_like(x)
        b[f] = bf
        b[p] = self.Afp.T @ x[f] + self.App @ xp
        self.x = x
        return x, b

    def sensitivity(self, dgdx=None, dgdb=None):
        f, p = self.f, self.p
        load = np.zeros_like(self.x[f])
        if dgdx is not None:
            load += dgdx[f]
        if dgdb is not None:
            load += self.Afp @ dgdb[p]
        lam = np.zeros_like(self.x)
        lam[f] = -self.lu.solve(load)
        if dgdb is not None:
            lam[p] = dgdb[p]
        dbf = -lam[f]
        dxp = self.Afp.T @ lam[f]
        if dgdx is not None:
            dxp = dxp + dgdx[p]
        if dgdb is not None:
            dbf = dbf + dgdb[f]
            dxp = dxp + self.App @ dgdb[p]
        return lam, self.x, dbf, dxp
