"""Multigrid-preconditioned CG on the device.

Mirrors ``pymoto/solvers/iterative.py``:

* ``DampedJacobi(A=None, w=1.0)`` - M = D/w (lines 21-47); on the device it is the smoother of
  the multigrid levels (fused with the operator in one kernel).
* ``GeometricMultigrid(domain, A=None, cycle='V', inner_level=None, smoother=None, smooth_steps=5)``
  (lines 124-256): trilinear transfers, Galerkin coarse operators R^T A R, V-cycle whose first
  pre-smoothing sweep starts from zero; nested instances form the hierarchy
  (``ex_compliance_multigrid.py:108-121``); the innermost level is solved directly.
* ``CG(A=None, preconditioner=..., tol=1e-7, maxit=10000, restart=50, verbosity=0)``
  (lines 295-403) with an explicit residual every ``restart`` iterations; a 2D right-hand side is
  solved by block PCG with the reference's Gram-Schmidt ``orth`` (lines 259-292), up to 8 columns
  per block, each iteration replayed from captured HIP graphs.

The hierarchy lives in one ``pmt_mg`` handle of the HIP runtime; ``CG.solve`` runs the whole
iteration inside the library (``pmt_pcg_solve``).
"""
import ctypes
import time
import warnings

import numpy as np
import torch

from .. import _lib
from ..domain import VoxelDomain
from .solvers import LinearSolver


class Preconditioner(LinearSolver):
    def update(self, A):
        pass

    def solve(self, rhs, x0=None, trans="N"):
        return _lib.as_device(rhs).clone()


class DampedJacobi(Preconditioner):
    """M = D / w; used as the multigrid smoother (``iterative.py:21-47``)."""

    def __init__(self, A=None, w=1.0):
        assert 0 < w <= 1, "w must be between 0 and 1"
        self.w = w
        self.D = None
        super().__init__(A)

    def update(self, A):
        self.D = A.diagonal()

    def solve(self, rhs, x0=None, trans="N"):
        """z = w r / diag(A) (``iterative.py:41-47``); a real matrix makes 'N', 'T' and 'H' the same."""
        if trans not in ("N", "T", "H"):
            raise TypeError("Only N, T, or H transposition is possible")
        if self.D is None:
            raise RuntimeError("DampedJacobi.update(A) must be called before solve")
        rhs = _lib.as_device(rhs)
        lib = _lib.lib()

        def one(c):
            z = torch.empty_like(c)
            _lib.check(lib.pmt_jacobi_apply(c.numel(), float(self.w), _lib.ptr(self.D), _lib.ptr(c), _lib.ptr(z),
                                            _lib.stream()), "Jacobi apply")
            return z
        if rhs.ndim == 1:
            return one(rhs)
        return torch.stack([one(rhs[:, k].contiguous()) for k in range(rhs.shape[1])], dim=1)


class ChebyshevSmoother:
    """Chebyshev polynomial smoother in D^-1 A for the device multigrid (no reference counterpart).

    Each level applies a polynomial of degree ``smooth_steps`` (as many operator sweeps as the damped
    Jacobi smoother it replaces) on the interval [lower*hi, hi]: on the fine grid hi is
    lambda_max(diag(K0)^-1 K0), an upper bound of lambda_max(D^-1 K) for any positive densities; on the
    Galerkin levels hi = upper * (power-iteration estimate).  Unlike a fixed omega it cannot exceed the
    Jacobi stability bound of a heterogeneous SIMP operator."""

    def __init__(self, lower: float = 0.25, upper: float = 1.1):
        self.lower, self.upper = float(lower), float(upper)

    def update(self, A):
        pass


class GeometricMultigrid(Preconditioner):
    """Geometric multigrid V-cycle preconditioner (``iterative.py:124-256``)."""

    _available_cycles = ["v", "w"]
    # V-cycle arithmetic: fp32 = FP32 throughout the V-cycle; bf16 = as fp32 with the fine-grid
    # operator on the matrix cores (BF16 operands, FP32 accumulation); CG stays FP64 in all cases
    _precision_bits = {"fp64": 64, "fp32": 32, "bf16": 16}

    def __init__(self, domain: VoxelDomain, A=None, cycle: str = "V", inner_level=None, smoother=None,
                 smooth_steps: int = 5, precision: str = "fp64"):
        self._handle = None
        assert domain.nelx % 2 == 0 and domain.nely % 2 == 0 and domain.nelz % 2 == 0, (
            f"Domain sizes {domain.nelx, domain.nely, domain.nelz} must be divisible by 2")
        assert cycle.lower() in self._available_cycles, f"Cycle ({cycle}) is not available"
        if cycle.lower() != "v":
            raise NotImplementedError("only the V-cycle is implemented (as in the reference)")
        self.domain = domain
        self.cycle = cycle
        self.inner_level = inner_level
        self.smoother = DampedJacobi(w=0.5) if smoother is None else smoother
        self.smooth_steps = smooth_steps
        if precision not in self._precision_bits:
            raise ValueError("precision must be 'fp64' (the reference's arithmetic), 'fp32' or 'bf16'")
        if precision == "bf16" and domain.dim != 3:
            raise NotImplementedError("the BF16 matrix-core operator is implemented for hex8 grids only")
        # arithmetic of the whole V-cycle (taken from the outermost level); not in the reference API
        self.precision = precision
        self.sub_domain = domain.coarsen()
        self.A = None
        self._handle = None
        self._key = None
        super().__init__(A)

    def levels(self):
        """The chain of GeometricMultigrid objects from this one to the innermost."""
        chain = [self]
        while isinstance(chain[-1].inner_level, GeometricMultigrid):
            nxt = chain[-1].inner_level
            sd = chain[-1].sub_domain
            if (nxt.domain.nelx, nxt.domain.nely, nxt.domain.nelz) != (sd.nelx, sd.nely, sd.nelz):
                raise ValueError("inner multigrid level does not match the coarse grid of its parent")
            chain.append(nxt)
        return chain

    @property
    def nlevels(self):
        return len(self.levels()) + 1

    def _ensure_handle(self, A):
        self._scalar = bool(getattr(A.grid, "scalar", False))
        chain = self.levels()
        omega = np.array([getattr(m.smoother, "w", 0.5) for m in chain], dtype=np.float64)
        cheb = isinstance(self.smoother, ChebyshevSmoother)
        steps = np.array([m.smooth_steps for m in chain], dtype=np.int32)
        d = self.domain
        key = (d.nelx, d.nely, d.nelz, len(chain) + 1, tuple(omega), tuple(steps), A.grid.K0.tobytes(),
               self.precision, (self.smoother.lower, self.smoother.upper) if cheb else None)
        if self._handle is not None and self._key == key:
            return
        self._release()
        lib = _lib.lib()
        h = ctypes.c_void_p()
        K0 = np.ascontiguousarray(A.grid.K0, dtype=np.float64)
        if self._scalar:
            if self.precision != "fp64" or cheb:
                raise NotImplementedError("scalar problems use the FP64 damped-Jacobi multigrid")
            _lib.check(lib.pmt_smg_create(d.nelx, d.nely, d.nelz if d.dim == 3 else 0, len(chain) + 1,
                                          omega.ctypes.data_as(ctypes.c_void_p), steps.ctypes.data_as(ctypes.c_void_p),
                                          K0.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)),
                       "scalar multigrid setup")
            self._handle, self._key = h, key
            return
        _lib.check(lib.pmt_mg_create(d.nelx, d.nely, d.nelz if d.dim == 3 else 0, len(chain) + 1,
                                     omega.ctypes.data_as(ctypes.c_void_p), steps.ctypes.data_as(ctypes.c_void_p),
                                     K0.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)), "multigrid setup")
        self._handle, self._key = h, key
        if self.precision != "fp64":
            _lib.check(lib.pmt_mg_set_precision(h, self._precision_bits[self.precision]), "multigrid precision")
        if cheb:
            _lib.check(lib.pmt_mg_set_smoother(h, 1, self.smoother.lower, self.smoother.upper), "multigrid smoother")

    def _release(self):
        if self._handle is not None:
            try:
                lib = _lib.load()
                (lib.pmt_smg_destroy if getattr(self, "_scalar", False) else lib.pmt_mg_destroy)(self._handle)
            except Exception:
                pass
            self._handle = None

    @property
    def scalar(self):
        """True when the hierarchy is the scalar (one unknown per node) one."""
        return bool(getattr(self, "_scalar", False))

    def __del__(self):
        self._release()

    def update(self, A):
        self.A = A
        self._ensure_handle(A)
        g = A.grid
        upd = _lib.lib().pmt_smg_update if self.scalar else _lib.lib().pmt_mg_update
        _lib.check(upd(self._handle, _lib.ptr(A.xe), _lib.ptr(g.bcmask_dev), g.bcdiag, _lib.stream()),
                   "multigrid update")

    @property
    def handle(self):
        return self._handle

    def device_bytes(self):
        if self._handle is None or self.scalar:
            return 0
        return _lib.load().pmt_mg_device_bytes(self._handle)

    def solve(self, rhs, x0=None, trans="N"):
        if x0 is not None:
            raise NotImplementedError("the device V-cycle starts from zero, as CG calls it")
        rhs = _lib.as_device(rhs)
        lib = _lib.lib()

        vcycle = lib.pmt_smg_vcycle if self.scalar else lib.pmt_mg_vcycle

        def one(c):
            z = torch.empty_like(c)
            _lib.check(vcycle(self._handle, _lib.ptr(c), _lib.ptr(z), _lib.stream()), "V-cycle")
            return z
        if rhs.ndim == 1:
            return one(rhs)
        return torch.stack([one(rhs[:, k].contiguous()) for k in range(rhs.shape[1])], dim=1)

    def smoother_interval(self, level):
        """(lo, hi) of the Chebyshev smoother on a level (after update)."""
        lo, hi = ctypes.c_double(0.0), ctypes.c_double(0.0)
        _lib.check(_lib.load().pmt_mg_smoother_interval(self._handle, int(level), ctypes.byref(lo), ctypes.byref(hi)),
                   "smoother interval")
        return lo.value, hi.value

    def level_apply(self, level, u):
        """y = A_level u with the Galerkin operator of that level (level 0 = fine operator)."""
        u = _lib.as_device(u)
        y = torch.empty_like(u)
        fn = _lib.lib().pmt_smg_level_apply if self.scalar else _lib.lib().pmt_mg_level_apply
        _lib.check(fn(self._handle, int(level), _lib.ptr(u), _lib.ptr(y), _lib.stream()), "level apply")
        return y


def auto_levels(domain: VoxelDomain) -> int:
    """Grid count of the reference example's hierarchy (``ex_compliance_multigrid.py:108-121``): one
    multigrid level per halving, adding levels while the coarse grid's extents are all even and at
    least 8; the last coarse grid is solved directly.  Returns the number of grids (fine included)."""
    nel = [domain.nelx, domain.nely] + ([domain.nelz] if domain.dim == 3 else [])
    if any(n % 2 for n in nel):
        return 1
    sub = [n // 2 for n in nel]
    levels = 2  # fine grid + the first coarse grid
    while all(n % 2 == 0 for n in sub) and all(n >= 8 for n in sub):
        sub = [n // 2 for n in sub]
        levels += 1
    return levels


def build_hierarchy(domain: VoxelDomain, levels: int, omega=0.5, smooth_steps=5, precision="fp64"):
    """Nested GeometricMultigrid objects with ``levels`` grids (fine grid included); omega is the damped
    Jacobi factor, or "chebyshev" for the Chebyshev smoother on every level."""
    if levels < 2:
        raise ValueError("a multigrid hierarchy needs at least two grids")

    def smoother():
        return ChebyshevSmoother() if omega == "chebyshev" else DampedJacobi(w=omega)
    mgs = [GeometricMultigrid(domain, smoother=smoother(), smooth_steps=smooth_steps, precision=precision)]
    for _ in range(levels - 2):
        mgs.append(GeometricMultigrid(mgs[-1].sub_domain, smoother=smoother(), smooth_steps=smooth_steps))
        mgs[-2].inner_level = mgs[-1]
    return mgs[0]


class CG(LinearSolver):
    """Preconditioned conjugate gradients (``iterative.py:295-403``), run inside the HIP runtime."""

    max_block = 8  # right-hand sides per block solve (pmt_block_pcg_solve)

    def __init__(self, A=None, preconditioner=None, tol: float = 1e-7, maxit: int = 10000, restart: int = 50,
                 verbosity: int = 0, graph: bool = True):
        self.preconditioner = preconditioner
        # several right-hand sides: replay each block-CG iteration as captured HIP graphs
        self.graph = graph
        self.A = A
        self.tol = tol
        self.maxit = maxit
        self.restart = restart
        self.verbosity = verbosity
        self.iterations = 0
        self.history = np.zeros(0)
        self.relres = np.nan
        super().__init__(A)

    def update(self, A):
        t0 = time.perf_counter()
        self.A = A
        if self.preconditioner is None:
            self.preconditioner = build_hierarchy(A.grid.domain, max(2, auto_levels(A.grid.domain)))
        self.preconditioner.update(A)
        if self.verbosity >= 1:
            torch.cuda.synchronize()
            print(f"CG Preconditioner set up in {np.round(time.perf_counter() - t0, 3)}s")

    def _device_mg(self):
        return isinstance(self.preconditioner, GeometricMultigrid)

    def _solve_generic(self, b, x0):
        """CG.solve for one right-hand side (``iterative.py:340-403``) with any preconditioner that has
        ``solve(r)`` (e.g. DampedJacobi): the host drives the iteration, every product, dot and update
        runs in the HIP library."""
        lib = _lib.lib()
        A, M = self.A, self.preconditioner
        n = b.numel()
        work = _lib.reduce_work()
        sc = torch.empty(2, dtype=torch.float64, device=b.device)

        def dot(u, v):
            _lib.check(lib.pmt_dot(n, _lib.ptr(u), _lib.ptr(v), _lib.ptr(sc), _lib.ptr(work), _lib.stream()), "dot")
            return float(sc[0].item())

        def axpby(a, u, bb, v):
            _lib.check(lib.pmt_axpby(n, float(a), _lib.ptr(u), float(bb), _lib.ptr(v), _lib.stream()), "axpby")

        x = torch.zeros_like(b) if x0 is None else _lib.as_device(x0).clone()
        r = A.residual(b, x)
        bn = np.sqrt(dot(b, b))
        rr = dot(r, r)
        if not np.isfinite(bn) or not np.isfinite(rr):
            raise _lib.NumericalError("PCG solve: non-finite right-hand side or start vector")
        hist = []
        if bn == 0.0:
            self.iterations, self.relres, self.history = 0, 0.0, np.zeros(1)
            return torch.zeros_like(b)
        tval = np.sqrt(rr) / bn
        hist.append(tval)
        i = -1
        if tval > self.tol:
            z = M.solve(r)
            p = torch.empty_like(z)
            axpby(1.0 / np.sqrt(dot(z, z)), z, 0.0, p)  # orth normalises the first direction
            for i in range(self.maxit):
                q = A @ p
                pq = dot(p, q)
                alpha = dot(p, r) / pq
                axpby(alpha, p, 1.0, x)
                if i % self.restart == 0:
                    r = A.residual(b, x)
                else:
                    axpby(-alpha, q, 1.0, r)
                tval = np.sqrt(dot(r, r)) / bn
                hist.append(tval)
                if not np.isfinite(tval):
                    raise _lib.NumericalError(f"PCG solve: non-finite residual at iteration {i + 1}")
                if tval <= self.tol:
                    break
                z = M.solve(r)
                beta = -dot(q, z) / pq
                axpby(1.0, z, beta, p)
        self.iterations, self.relres, self.history = i + 1, tval, np.asarray(hist)
        if tval > self.tol:
            warnings.warn(f"CG Maximum iterations ({self.maxit}) reached, with final residuals {tval}")
        return x

    def _solve_1rhs(self, b, x0, x0_scratch=False):
        """x0_scratch: x0 is a device vector the caller does not need again, so the solve may iterate in it
        instead of a copy."""
        if not self._device_mg():
            return self._solve_generic(b, x0)
        lib = _lib.lib()
        if x0 is None:
            x = torch.zeros_like(b)
        elif x0_scratch and isinstance(x0, torch.Tensor) and x0.is_cuda and x0.dtype == torch.float64 \
                and x0.is_contiguous() and x0.shape == b.shape:
            x = x0
        else:
            x = _lib.as_device(x0).clone()
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        hist = np.zeros(self.maxit + 1)
        pcg = lib.pmt_smg_pcg_solve if self.preconditioner.scalar else lib.pmt_pcg_solve
        _lib.check(pcg(self.preconditioner.handle, _lib.ptr(b), _lib.ptr(x), int(x0 is None), float(self.tol),
                       int(self.maxit), int(self.restart), ctypes.byref(it), ctypes.byref(rel),
                       hist.ctypes.data_as(ctypes.c_void_p), _lib.stream()), "PCG solve")
        self.iterations, self.relres = it.value, rel.value
        self.history = hist[: it.value + 1]
        if rel.value > self.tol:
            warnings.warn(f"CG Maximum iterations ({self.maxit}) reached, with final residuals {rel.value}")
        elif self.verbosity >= 1:
            print(f"CG Converged in {it.value} iterations, with final residual {rel.value}")
        return x

    def _solve_block(self, B, X0):
        """Block PCG (``iterative.py:340-403`` with a 2D right-hand side) for 2..8 columns.  B, X0: (k, n)."""
        lib = _lib.lib()
        k = B.shape[0]
        X = torch.zeros_like(B) if X0 is None else X0.clone()
        it = ctypes.c_int(0)
        rel = np.zeros(k)
        hist = np.zeros(self.maxit + 1)
        _lib.check(lib.pmt_block_pcg_solve(self.preconditioner.handle, k, _lib.ptr(B), _lib.ptr(X), int(X0 is None),
                                           float(self.tol), int(self.maxit), int(self.restart), int(bool(self.graph)),
                                           ctypes.byref(it), rel.ctypes.data_as(ctypes.c_void_p),
                                           hist.ctypes.data_as(ctypes.c_void_p), _lib.stream()), "block PCG solve")
        self.iterations, self.relres = it.value, rel
        self.history = hist[: it.value + 1]
        if rel.max() > self.tol:
            warnings.warn(f"CG Maximum iterations ({self.maxit}) reached, with final residuals {rel}")
        elif self.verbosity >= 1:
            print(f"CG Converged in {it.value} iterations, with final (max) residual {rel.max()}")
        return X

    def solve(self, rhs, x0=None, trans="N"):
        rhs = _lib.as_device(rhs)
        if rhs.ndim == 1:
            return self._solve_1rhs(rhs, x0)
        if rhs.ndim == 2 and rhs.shape[1] == 1:
            return self._solve_1rhs(rhs[:, 0].contiguous(), None if x0 is None else _lib.as_device(x0).reshape(-1)
                                    ).reshape(-1, 1)
        if not self._device_mg() or self.preconditioner.scalar:
            # column by column (block PCG runs on the elasticity multigrid handle)
            X0 = None if x0 is None else _lib.as_device(x0)
            cols, its = [], []
            for k in range(rhs.shape[1]):
                cols.append(self._solve_1rhs(rhs[:, k].contiguous(), None if X0 is None else X0[:, k].contiguous()))
                its.append(self.iterations)
            self.iterations = max(its)
            return torch.stack(cols, dim=1)
        Bt = rhs.t().contiguous()  # column j contiguous
        X0t = None if x0 is None else _lib.as_device(x0).t().contiguous()
        out, its = [], []
        for c0 in range(0, Bt.shape[0], self.max_block):
            blk = slice(c0, c0 + self.max_block)
            out.append(self._solve_block(Bt[blk].contiguous(), None if X0t is None else X0t[blk].contiguous()))
            its.append(self.iterations)
        self.iterations = max(its)
        return torch.cat(out, dim=0).t().contiguous()


def auto_determine_solver(A, tol=1e-12, **kwargs):
    """Default solver for a device StiffnessMatrix: MG-preconditioned CG to a tight tolerance, in place of
    the reference's sparse direct solve (``pymoto/solvers/auto_determine.py:109-122``)."""
    return CG(preconditioner=build_hierarchy(A.grid.domain, max(2, auto_levels(A.grid.domain))), tol=tol)
