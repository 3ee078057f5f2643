"""Oracle: DampedJacobi, GeometricMultigrid, CG, LDAS and the direct solve (test infrastructure only).

Restates ``pymoto/solvers/iterative.py:21-47`` (DampedJacobi), ``124-256`` (GeometricMultigrid:
trilinear R with weights 1/.5/.25/.125, Galerkin ``R^T A R``, V-cycle with the first pre-smoothing
sweep from zero), ``259-292`` (orth), ``295-403`` (block PCG with explicit residual every
``restart`` iterations including i=0), ``solvers.py:88-306`` (LDAWrapper) and
``sparse.py:522-550`` (SuperLU direct solve, the choice of ``auto_determine_solver`` without MKL).

Pinning (tests/test_oracle_solvers.py): R^T 1 = 4/3/2.25 (2D) and 8/6/4.5/3.375 (3D), R 1 = 1
(``tests/test_solvers_multigrid.py:9-91`` of the reference); CG and MG-CG solutions against
``scipy.sparse.linalg.spsolve`` and the residual ||Ku-f||/||f||; CG iteration counts and final
residuals against golden runs of the reference's own GMG+CG.
"""
import warnings

import numpy as np
import scipy.sparse as sps
from scipy.sparse.linalg import splu


class DampedJacobi:
    def __init__(self, w=1.0):
        assert 0 < w <= 1
        self.w = w
        self.D = None

    def update(self, A):
        self.D = A.diagonal()

    def solve(self, rhs, x0=None):
        return self.w * (rhs.T / self.D).T


class DirectLU:
    def update(self, A):
        self.lu = splu(sps.csc_matrix(A))

    def solve(self, rhs, x0=None):
        return self.lu.solve(rhs)


def interpolation(nelx, nely, nelz, ndof):
    """Prolongation R (n_fine x n_coarse), iterative.py:178-220."""
    cx, cy, cz = nelx // 2, nely // 2, nelz // 2
    dim = 3 if nelz > 0 else 2
    w = np.ones((3, 3, 3)) * 0.125
    w[1, :, :] = 0.25
    w[:, 1, :] = 0.25
    w[:, :, 1] = 0.25
    w[1, 1, :] = 0.5
    w[1, :, 1] = 0.5
    w[:, 1, 1] = 0.5
    w[1, 1, 1] = 1.0
    rows, cols, vals = [], [], []

    def cnode(i, j, k):
        return (k * (cy + 1) + j) * (cx + 1) + i

    def fnode(i, j, k):
        return (k * (nely + 1) + j) * (nelx + 1) + i

    for i in (-1, 0, 1):
        ix = np.arange(max(-i, 0), min(cx + 1 - i, cx + 1))
        for j in (-1, 0, 1):
            iy = np.arange(max(-j, 0), min(cy + 1 - j, cy + 1))
            for k in ((-1, 0, 1) if dim == 3 else (0,)):
                iz = np.arange(max(-k, 0), min(cz + 1 - k, cz + 1))
                gi, gj, gk = np.meshgrid(ix, iy, iz, indexing="ij")
                nc = cnode(gi, gj, gk).ravel()
                nf = fnode(2 * gi + i, 2 * gj + j, 2 * gk + k).ravel()
                for d in range(ndof):
                    rows.append(nf * ndof + d)
                    cols.append(nc * ndof + d)
                    vals.append(np.full(nf.size, w[1 + i, 1 + j, 1 + k]))
    nf_tot = ndof * (nelx + 1) * (nely + 1) * (nelz + 1)
    nc_tot = ndof * (cx + 1) * (cy + 1) * (cz + 1)
    R = sps.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(nf_tot, nc_tot))
    return sps.csr_matrix(R)


class GeometricMultigrid:
    def __init__(self, nel, inner=None, smoother=None, smooth_steps=5):
        nelx, nely, nelz = nel
        assert nelx % 2 == 0 and nely % 2 == 0 and nelz % 2 == 0
        self.nel = (nelx, nely, nelz)
        self.sub_nel = (nelx // 2, nely // 2, nelz // 2)
        self.inner = inner
        self.smoother = DampedJacobi(w=0.5) if smoother is None else smoother
        self.smooth_steps = smooth_steps
        self.R = None

    def update(self, A):
        if self.R is None:
            n_nodes = (self.nel[0] + 1) * (self.nel[1] + 1) * (self.nel[2] + 1)
            self.R = interpolation(*self.nel, A.shape[0] // n_nodes)
        self.A = A
        self.smoother.update(A)
        self.Ac = (self.R.T @ A @ self.R).tocsr()
        if self.inner is None:
            self.inner = DirectLU()
        self.inner.update(self.Ac)

    def solve(self, rhs, x0=None):
        A = self.A
        if x0 is None:
            u = self.smoother.solve(rhs)
        else:
            u = x0 + self.smoother.solve(rhs - A @ x0)
        for _ in range(self.smooth_steps - 1):
            u += self.smoother.solve(rhs - A @ u)
        r = rhs - A @ u
        uc = self.inner.solve(self.R.T @ r)
        u += self.R @ uc
        for _ in range(self.smooth_steps):
            u += self.smoother.solve(rhs - A @ u)
        return u


def build_mg(nel, levels, omega=0.5, smooth_steps=5):
    """Nested GMG hierarchy as in ex_compliance_multigrid.py:108-121; ``levels`` counts the fine grid."""
    mgs = [GeometricMultigrid(nel, smoother=DampedJacobi(omega), smooth_steps=smooth_steps)]
    for _ in range(levels - 2):
        sub = mgs[-1].sub_nel
        mgs.append(GeometricMultigrid(sub, smoother=DampedJacobi(omega), smooth_steps=smooth_steps))
        mgs[-2].inner = mgs[-1]
    return mgs[0]


def orth(u, normalize=True, zero_rtol=1e-15):
    if u.ndim == 1:
        return u
    vecs = []
    for i in range(u.shape[-1]):
        vi = np.copy(u[..., i])
        b0 = vi @ vi
        for vj in vecs:
            vi -= vj * (vi @ vj) / (1.0 if normalize else vj @ vj)
        b1 = vi @ vi
        if b1 / b0 < zero_rtol:
            continue
        if normalize:
            vi /= np.sqrt(b1)
        vecs.append(vi)
    return np.stack(vecs, axis=-1)


class CG:
    def __init__(self, preconditioner=None, tol=1e-7, maxit=10000, restart=50):
        self.M = preconditioner
        self.tol, self.maxit, self.restart = tol, maxit, restart
        self.iterations = 0
        self.history = []

    def update(self, A):
        self.A = A
        if self.M is not None:
            self.M.update(A)

    def _prec(self, r):
        return r.copy() if self.M is None else self.M.solve(r)

    def solve(self, rhs, x0=None):
        A = self.A
        b = rhs.reshape(-1, 1) if rhs.ndim == 1 else rhs
        x = np.zeros_like(b, dtype=float) if x0 is None else x0.reshape(b.shape).astype(float).copy()
        r = b - A @ x
        bn = np.linalg.norm(b, axis=0)
        tval = np.linalg.norm(r, axis=0) / bn
        self.history = [tval.max()]
        self.iterations = 0
        if tval.max() <= self.tol:
            return x.ravel() if rhs.ndim == 1 else x
        z = self._prec(r)
        p = orth(z, normalize=True)
        i = 0
        for i in range(self.maxit):
            q = A @ p
            pq = p.T @ q
            pq_inv = np.linalg.inv(pq)
            alpha = pq_inv @ (p.T @ r)
            x += p @ alpha
            if i % self.restart == 0:
                r = b - A @ x
            else:
                r -= q @ alpha
            tval = np.linalg.norm(r, axis=0) / bn
            self.history.append(tval.max())
            if tval.max() <= self.tol:
                break
            z = self._prec(r)
            beta = -pq_inv @ (q.T @ z)
            p = orth(z + p @ beta, normalize=False)
        self.iterations = i + 1
        if tval.max() > self.tol:
            warnings.warn(f"CG Maximum iterations ({self.maxit}) reached, with final residuals {tval}")
        return x.ravel() if rhs.ndim == 1 else x


def diagonal_rows(A):
    """Rows whose only nonzero is the diagonal (solvers.py:88-96, including its logical_and quirk)."""
    bmat = A != 0
    has_diag = bmat.diagonal()
    nnz_rows = np.array(bmat.sum(axis=0)).ravel()
    return np.logical_and(has_diag, nnz_rows <= 1)


class LDAS:
    """Linear-dependency-aware wrapper for symmetric matrices (solvers.py:99-306, normal storage)."""

    def __init__(self, solver, tol=1e-7):
        self.solver, self.tol = solver, tol
        self.xs, self.bs = [], []
        self.did_solve = False

    def update(self, A):
        self.A = A
        d = diagonal_rows(A)
        self.idia = np.flatnonzero(d)
        self.isel = np.flatnonzero(~d)
        self.xs.clear()
        self.bs.clear()
        self.solver.update(A)

    def solve(self, rhs, x0=None):
        A, isel, idia = self.A, self.isel, self.idia
        rl = rhs.reshape(-1, 1).astype(float).copy() if rhs.ndim == 1 else rhs.astype(float).copy()
        sol = np.zeros_like(rl)
        sol[idia] = rl[idia] / A.diagonal()[idia, None]
        rl[idia] = 0
        for x, b in zip(self.xs, self.bs):
            alpha = rl[isel].T @ b / (b @ b)
            rl[isel] -= alpha * b[:, None]
            sol[isel] += alpha * x[:, None]
        ref = rhs if rhs.ndim > 1 else rhs.reshape(-1, 1)
        bn = np.linalg.norm(ref, axis=0)
        bn[bn == 0] = 1
        rtol = np.linalg.norm(A @ sol - ref, axis=0) / bn
        self.did_solve = rtol > self.tol
        if np.any(self.did_solve):
            if x0 is not None:
                x0l = (x0.reshape(-1, 1) if x0.ndim == 1 else x0[..., self.did_solve]).astype(float).copy()
                x0l[idia] = 0
                for x in self.xs:
                    x0l[isel] -= (x0l[isel].T @ x / (x @ x)) * x[:, None]
            else:
                x0l = None
            xnew = self.solver.solve(rl[:, self.did_solve], x0l)
            xnew = xnew.reshape(-1, int(np.sum(self.did_solve)))
            sol[np.ix_(isel, self.did_solve)] += xnew[isel]
            for i in range(xnew.shape[-1]):
                xadd = xnew[isel, i].copy()
                badd = (A @ xnew[:, i])[isel]
                for x, b in zip(self.xs, self.bs):
                    beta = badd @ b / (b @ b)
                    badd -= beta * b
                    xadd -= beta * x
                bn1 = np.linalg.norm(badd)
                if not np.isfinite(bn1) or bn1 == 0:
                    continue
                self.xs.append(xadd / bn1)
                self.bs.append(badd / bn1)
        return sol.ravel() if rhs.ndim == 1 else sol
