"""Failure handling of the device solves, the coarsest-inverse join across streams, and the
reference-side ctypes binding.

Every test here ends with every kernel exiting normally: non-finite values are ordinary IEEE data
to the kernels (no kernel indexes memory with a solution value), and the host turns them into an
error status after one iteration.
"""
import time

import numpy as np
import numpy.testing as npt
import pytest
import torch

import pymoto_amd as pm
from pymoto_amd import _lib
from oracle import fe, problems

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64).cuda()


def host(t):
    return t.detach().cpu().numpy()


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300))


def _system(nel=(16, 8, 8), seed=0):
    rng = np.random.default_rng(seed)
    d = pm.VoxelDomain(*nel)
    bc, f = problems.cantilever3d(*nel)
    K = pm.AssembleStiffness(d, bc=bc)(dev(fe.simp(rng.uniform(0.05, 1.0, d.nel))))
    return d, bc, f, K


def _cg(d, levels=3, **kw):
    return pm.solvers.CG(preconditioner=pm.solvers.build_hierarchy(d, levels, omega=0.5, smooth_steps=2), **kw)


def test_cg_warns_at_maxit(gpu_lib):
    """CG.solve warns when maxit ends the iteration above tol (iterative.py:398-401) and returns the iterate."""
    d, _, f, K = _system()
    cg = _cg(d, tol=1e-14, maxit=2)
    cg.update(K)
    with pytest.warns(UserWarning, match="Maximum iterations"):
        u = cg.solve(dev(f))
    assert cg.iterations == 2 and np.isfinite(cg.relres) and cg.relres > 1e-14
    assert np.all(np.isfinite(host(u)))


def test_block_cg_warns_at_maxit(gpu_lib):
    d, _, f, K = _system()
    F = np.stack([f, np.roll(f, 3)], axis=1)
    cg = _cg(d, tol=1e-14, maxit=2)
    cg.update(K)
    with pytest.warns(UserWarning, match="Maximum iterations"):
        cg.solve(dev(F))
    assert cg.iterations == 2


@pytest.mark.parametrize("where", ["rhs", "x0"])
def test_cg_non_finite_input_raises_at_once(gpu_lib, where):
    """A NaN in the right-hand side or the start vector ends the solve with PMT_ERR_NUMERIC at once
    (no polling timeout per iteration, no maxit iterations)."""
    d, _, f, K = _system()
    cg = _cg(d, tol=1e-10, maxit=3)
    cg.update(K)
    b = f.copy()
    x0 = None
    if where == "rhs":
        b[5] = np.nan
    else:
        x0 = dev(np.where(np.arange(f.size) == 7, np.inf, 0.0))
    t0 = time.perf_counter()
    with pytest.raises(_lib.NumericalError, match="non-finite"):
        cg.solve(dev(b), x0=x0)
    assert time.perf_counter() - t0 < 0.9  # the stopping-test poll never waits out its 1 s fallback
    # the handle still works afterwards
    cg.maxit = 500
    u = cg.solve(dev(f))
    assert cg.relres <= 1e-10 and np.all(np.isfinite(host(u)))


def test_cg_non_finite_operator_raises_in_the_loop(gpu_lib):
    """A NaN element factor (finite right-hand side, zero start) first shows in r.r after the first
    iteration: the polled stopping test reads it (sequence-numbered publish, not a NaN marker) and
    the solve ends with PMT_ERR_NUMERIC after one iteration.  Every kernel treats the NaN as data
    (no density is used as an index or a loop bound), so all of them exit normally."""
    rng = np.random.default_rng(3)
    nel = (16, 8, 8)
    d = pm.VoxelDomain(*nel)
    bc, f = problems.cantilever3d(*nel)
    x = fe.simp(rng.uniform(0.05, 1.0, d.nel))
    x[d.nel // 2] = np.nan
    K = pm.AssembleStiffness(d, bc=bc)(dev(x))
    cg = _cg(d, tol=1e-10, maxit=50)
    cg.update(K)
    t0 = time.perf_counter()
    with pytest.raises(_lib.NumericalError, match="iteration 1 "):
        cg.solve(dev(f))
    assert time.perf_counter() - t0 < 0.9


def test_block_cg_non_finite_rhs_raises(gpu_lib):
    d, _, f, K = _system()
    F = np.stack([f, np.roll(f, 3)], axis=1)
    F[11, 1] = np.nan
    cg = _cg(d, tol=1e-10, maxit=3)
    cg.update(K)
    with pytest.raises(_lib.NumericalError, match="non-finite"):
        cg.solve(dev(F))


def test_cg_zero_rhs_returns_zero(gpu_lib):
    d, _, f, K = _system()
    cg = _cg(d, tol=1e-10)
    cg.update(K)
    u = cg.solve(dev(np.zeros_like(f)), x0=dev(np.ones_like(f)))
    assert cg.iterations == 0 and cg.relres == 0.0
    assert not np.any(host(u))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_coarsest_inverse_join_each_consumer_stream(gpu_lib, prec):
    """Update on stream A (the coarsest inversion forked to the handle's own stream), a V-cycle on stream
    B ordered after A's update, then a V-cycle on A ordered after B's (calls on one handle share its
    work vectors, so they must not overlap).  Each stream joins the inversion before it reads the
    inverse; results equal a hierarchy updated and solved on one stream, bit for bit."""
    rng = np.random.default_rng(11)
    nel = (32, 16, 16)
    d = pm.VoxelDomain(*nel)
    bc = problems.cantilever3d(*nel)[0]
    K = pm.AssembleStiffness(d, bc=bc)(dev(fe.simp(rng.uniform(0.05, 1.0, d.nel))))
    r = dev(rng.standard_normal(K.shape[0]))
    ref = pm.solvers.build_hierarchy(d, 4, omega=0.5, smooth_steps=2, precision=prec)
    ref.update(K)
    want = host(ref.solve(r))
    mg = pm.solvers.build_hierarchy(d, 4, omega=0.5, smooth_steps=2, precision=prec)
    A, B = torch.cuda.Stream(), torch.cuda.Stream()
    A.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(A):
        mg.update(K)
    B.wait_stream(A)
    with torch.cuda.stream(B):
        zb = mg.solve(r)
    A.wait_stream(B)
    with torch.cuda.stream(A):
        za = mg.solve(r)
    torch.cuda.synchronize()
    npt.assert_array_equal(host(za), want)
    npt.assert_array_equal(host(zb), want)


def test_reference_side_ctypes_binding_two_designs(gpu_lib):
    """The torch-free binding a pyMOTO maintainer would add (pymoto_amd.ref_binding.SolverMGPCG) driven
    through two design updates with a warm start, as LinSolve drives its solver, against the device
    LinSolve on the same designs."""
    from pymoto_amd.ref_binding import SolverMGPCG

    class DesignSignal:  # the pyMOTO Signal feeding AssembleStiffness (only .state is read)
        state = None

    nel = (32, 16, 16)
    d = pm.VoxelDomain(*nel)
    bc, f = problems.cantilever3d(*nel)
    rng = np.random.default_rng(21)
    designs = [fe.simp(rng.uniform(0.05, 1.0, d.nel)) for _ in range(2)]
    K0 = fe.element_stiffness(3)
    sig = DesignSignal()
    stub = SolverMGPCG(d, K0, bc, float(np.max(K0)), sig, levels=4, omega=0.5, steps=2, tol=1e-10)
    sx = pm.Signal("xs", dev(designs[0]))
    lin = pm.LinSolve(hermitian=True, solver=pm.solvers.CG(
        preconditioner=pm.solvers.build_hierarchy(d, 4, omega=0.5, smooth_steps=2), tol=1e-10))
    with pm.Network() as fn:
        sK = pm.AssembleStiffness(d, bc=bc)(sx)
        su = lin(sK, dev(f))
    u_stub = None
    try:
        for x in designs:
            sig.state = x
            stub.update(None)
            u_stub = stub.solve(f, x0=u_stub)
            assert stub.relres <= 1e-10 and stub.iterations > 0
            sx.state = dev(x)
            fn.response()
            u_dev = host(su.state)
            assert relerr(u_stub, u_dev) < 1e-8
            Kx = fe.Assembly(*nel, K0, bc).matrix(x)
            assert np.linalg.norm(Kx @ u_stub - f) / np.linalg.norm(f) < 1e-9
        # a transposed solve is the same solve (symmetric K); a 2D right-hand side goes column by column
        U = stub.solve(np.stack([f, 2.0 * f], axis=1), trans="T")
        assert relerr(U[:, 1], 2.0 * U[:, 0]) < 1e-8
    finally:
        stub.close()


def test_stopping_test_never_reads_another_handles_publish(gpu_lib):
    """The polled stopping test of pmt_pcg_solve matches a sequence number in pinned memory.  A handle
    that published a NaN r.r and was destroyed leaves that memory (value and sequence number) to the next
    handle's pinned allocation; sequence numbers unique in the process (and a cleared slot) keep the next
    solve from taking the stale NaN for its own r.r.  Alternating NaN solves and normal solves on fresh
    handles must give finite, converged results every time."""
    rng = np.random.default_rng(17)
    nel = (16, 8, 8)
    d = pm.VoxelDomain(*nel)
    bc, f = problems.cantilever3d(*nel)
    x = fe.simp(rng.uniform(0.05, 1.0, d.nel))
    xbad = x.copy()
    xbad[d.nel // 3] = np.nan
    for trial in range(12):
        K = pm.AssembleStiffness(d, bc=bc)(dev(xbad))
        cg = _cg(d, tol=1e-10, maxit=20)
        cg.update(K)
        with pytest.raises(_lib.NumericalError):
            cg.solve(dev(f))
        del cg, K
        torch.cuda.synchronize()
        K = pm.AssembleStiffness(d, bc=bc)(dev(x))
        cg = _cg(d, tol=1e-10, maxit=200)
        cg.update(K)
        u = cg.solve(dev(f))
        assert cg.relres <= 1e-10 and np.all(np.isfinite(host(u))), trial
