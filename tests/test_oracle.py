"""Pin the CPU oracle against the reference's golden vectors, its own known answers and closed forms."""
import hashlib

import numpy as np
import numpy.testing as npt
import pytest
import scipy.sparse as sps
from scipy.sparse.linalg import spsolve

from conftest import golden
from oracle import fe, problems
from oracle.filter import ConvFilter, filter_weights
from oracle.pipeline import design_iteration
from oracle.solvers import CG, build_mg, interpolation


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---------------------------------------------------------------- element stiffness
def test_k0_matches_reference_golden():
    g = golden("k0_cases.npz")
    for i in range(int(g["ncases"])):
        is3d, sx, sy, sz = g[f"spec_{i}"]
        K0 = fe.element_stiffness(3 if is3d else 2, (sx, sy, sz), E=2.5, nu=0.27, plane=str(g[f"plane_{i}"]))
        npt.assert_allclose(K0, g[f"K0_{i}"], rtol=1e-13, atol=1e-13 * np.abs(K0).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_k0_rigid_body_nullspace_and_symmetry(dim):
    K0 = fe.element_stiffness(dim, (0.7, 1.3, 0.9))
    npt.assert_allclose(K0, K0.T, atol=1e-14)
    ev = np.linalg.eigvalsh(K0)
    nrbm = 3 if dim == 2 else 6
    assert np.all(np.abs(ev[:nrbm]) < 1e-12 * ev[-1]) and ev[nrbm] > 1e-6 * ev[-1]


def test_one_element_tensile_2d_closed_form():
    """u = F L /(E A) and Poisson contraction (reference tests/test_assembly.py:101-153)."""
    Lx, Ly, Lz, E, nu, F = 0.1, 0.2, 0.3, 210e9, 0.3, 1.5
    K0 = fe.element_stiffness(2, (Lx, Ly, Lz), E=E, nu=nu, plane="stress")
    asm = fe.Assembly(1, 1, 0, K0, bc=np.array([0, 4, 1]))
    K = asm.matrix(np.ones(1)).toarray()
    f = np.zeros(8)
    f[[2, 6]] = F / 2
    u = np.linalg.solve(K, f)
    ux = F * Lx / (E * Ly * Lz)
    npt.assert_allclose(u[[2, 6]], ux, rtol=1e-10)
    npt.assert_allclose(u[7], -nu * ux / Lx * Ly, rtol=1e-10)
    npt.assert_allclose(u[3], 0, atol=1e-10)


def test_one_element_tensile_3d_closed_form():
    """reference tests/test_assembly.py:156-207."""
    Lx, Ly, Lz, E, nu, F = 0.1, 0.2, 0.3, 210e9, 0.3, 1.5e3
    K0 = fe.element_stiffness(3, (Lx, Ly, Lz), E=E, nu=nu)
    n = lambda i, j, k: (k * 2 + j) * 2 + i  # noqa: E731
    left = [n(0, j, k) for j in range(2) for k in range(2)]
    right = [n(1, j, k) for j in range(2) for k in range(2)]
    bc = np.concatenate([np.array(left) * 3, np.array([n(0, 0, 0), n(0, 0, 1)]) * 3 + 1,
                         np.array([n(0, 0, 0), n(0, 1, 0)]) * 3 + 2])
    K = fe.Assembly(1, 1, 1, K0, bc=bc).matrix(np.ones(1)).toarray()
    f = np.zeros(24)
    f[np.array(right) * 3] = F / 4
    u = np.linalg.solve(K, f)
    ux = F * Lx / (E * Ly * Lz)
    npt.assert_allclose(u[np.array(right) * 3], ux, rtol=1e-10)
    npt.assert_allclose(u[n(1, 1, 0) * 3 + 1], -nu * ux / Lx * Ly, rtol=1e-10)
    npt.assert_allclose(u[n(1, 0, 1) * 3 + 2], -nu * ux / Lx * Lz, rtol=1e-10)


# ---------------------------------------------------------------- CSR pattern and values
def test_csr_pattern_and_values_bit_exact_vs_reference():
    g = golden("csr_cases.npz")
    for i in range(int(g["ncases"])):
        nel = tuple(int(v) for v in g[f"nel_{i}"])
        dim = 3 if nel[2] > 0 else 2
        K0 = fe.element_stiffness(dim, plane=str(g[f"plane_{i}"]))
        asm = fe.Assembly(*nel, K0, bc=g[f"bc_{i}"])
        npt.assert_array_equal(asm.indptr, g[f"indptr_{i}"].astype(np.int64))
        npt.assert_array_equal(asm.indices, g[f"indices_{i}"].astype(np.int64))
        data = asm.values(g[f"x_{i}"])
        assert data.tobytes() == g[f"data_{i}"].tobytes()


def test_csr_c1_hashes():
    g = golden("c1_mbb.npz")
    K0 = fe.element_stiffness(2, plane="stress")
    npt.assert_array_equal(K0, g["K0"])
    asm = fe.Assembly(120, 40, 0, K0, bc=g["bc"])
    assert sha(asm.indptr) == str(g["sha_indptr"])
    assert sha(asm.indices) == str(g["sha_indices"])
    assert sha(asm.values(g["xsimp"])) == str(g["sha_data"])
    assert str(g["index_dtype"]) == "int32"


def test_csr_brute_force_dense():
    nel = (3, 2, 2)
    K0 = fe.element_stiffness(3)
    x = np.random.default_rng(0).random(12)
    asm = fe.Assembly(*nel, K0)
    dc = fe.dofconn(*nel)
    Kd = np.zeros((asm.n, asm.n))
    for e in range(12):
        Kd[np.ix_(dc[e], dc[e])] += x[e] * K0
    npt.assert_allclose(asm.matrix(x).toarray(), Kd, rtol=1e-15, atol=1e-15)
    coo = sps.coo_matrix((np.ones(dc.size * 24), (np.repeat(dc, 24, axis=1).ravel(), np.tile(dc, (1, 24)).ravel())),
                         shape=(asm.n, asm.n)).tocsr()
    coo.sum_duplicates()
    coo.sort_indices()
    npt.assert_array_equal(coo.indptr, asm.indptr)
    npt.assert_array_equal(coo.indices, asm.indices)


# ---------------------------------------------------------------- filter
def _bcs_from(g, i):
    raw = [str(s) for s in g[f"case{i}_bcs"]]
    conv = [s if s in ("symmetric", "edge", "wrap") else float(s) for s in raw]
    return ((conv[0], conv[1]), (conv[2], conv[3]), (conv[4], conv[5]))


def test_filter_vs_reference_golden():
    g = golden("filter_cases.npz")
    for i in range(int(g["ncases"])):
        nel = tuple(int(v) for v in g[f"case{i}_nel"])
        r = float(g[f"case{i}_radius"])
        w = g[f"case{i}_weights"]
        if np.isfinite(r):
            npt.assert_allclose(filter_weights(r, nel), w, rtol=1e-15)
        flt = ConvFilter(nel, w, _bcs_from(g, i))
        npt.assert_allclose(flt.forward(g[f"case{i}_x"]), g[f"case{i}_y"], rtol=1e-12, atol=1e-14)
        npt.assert_allclose(flt.adjoint(g[f"case{i}_dy"]), g[f"case{i}_dx"], rtol=1e-12, atol=1e-14)


def test_filter_known_answers_2d_interior_and_xmin_mirror():
    """reference tests/test_filter.py:13-66."""
    nel = (10, 12, 0)
    w = filter_weights(2.0, nel)
    e = lambda i, j: j * 10 + i  # noqa: E731
    x = np.zeros(120)
    x[e(5, 6)] = 1.0
    y = ConvFilter(nel, w).forward(x)
    npt.assert_allclose(y[e(5, 6)], w[1, 1, 0])
    npt.assert_allclose(y[e(6, 7)], w[0, 0, 0])
    x = np.zeros(120)
    x[e(0, 6)] = 1.0
    y = ConvFilter(nel, w).forward(x)
    npt.assert_allclose(y[e(0, 6)], w[1, 1, 0] + w[2, 1, 0])
    npt.assert_allclose(y[e(0, 5)], w[1, 2, 0] + w[2, 2, 0])
    y = ConvFilter(nel, w, ((0.0, "symmetric"), ("symmetric",) * 2, ("symmetric",) * 2)).forward(x)
    npt.assert_allclose(y[e(0, 6)], w[1, 1, 0])


def test_filter_adjoint_identity_all_bcs():
    rng = np.random.default_rng(3)
    nel = (7, 6, 5)
    w = rng.random((3, 5, 3))
    flt = ConvFilter(nel, w, (("edge", 0.3), ("wrap", "wrap"), (1.0, "symmetric")))
    # constant padding is affine: check adjointness of the linear part
    x, y = rng.random(210), rng.random(210)
    lin = flt.forward(x) - flt.forward(np.zeros(210))
    npt.assert_allclose(lin @ y, x @ flt.adjoint(y), rtol=1e-12)


# ---------------------------------------------------------------- multigrid transfers
def test_restriction_sums_2d_3d():
    """reference tests/test_solvers_multigrid.py:9-91."""
    R = interpolation(10, 10, 0, 2)
    uc = (R.T @ np.ones(R.shape[0])).reshape(6, 6, 2)
    npt.assert_allclose(uc[1:-1, 1:-1], 4)
    npt.assert_allclose(uc[0, 1:-1], 3)
    npt.assert_allclose(uc[0, 0], 2.25)
    npt.assert_allclose(R @ np.ones(R.shape[1]), 1.0)
    R = interpolation(10, 10, 10, 3)
    uc = (R.T @ np.ones(R.shape[0])).reshape(6, 6, 6, 3)
    npt.assert_allclose(uc[1:-1, 1:-1, 1:-1], 8)
    npt.assert_allclose(uc[0, 1:-1, 1:-1], 6)
    npt.assert_allclose(uc[0, 0, 1:-1], 4.5)
    npt.assert_allclose(uc[0, 0, 0], 3.375)
    npt.assert_allclose(R @ np.ones(R.shape[1]), 1.0)


def test_mgcg_matches_spsolve():
    nel = (16, 8, 8)
    bc, f = problems.cantilever3d(*nel)
    x = np.random.default_rng(0).uniform(0.1, 1.0, 16 * 64)
    asm = fe.Assembly(*nel, fe.element_stiffness(3), bc=bc)
    K = asm.matrix(fe.simp(x))
    cg = CG(preconditioner=build_mg(nel, 3, omega=0.6, smooth_steps=2), tol=1e-12)
    cg.update(K)
    u = cg.solve(f)
    npt.assert_allclose(u, spsolve(K.tocsc(), f), rtol=1e-8, atol=1e-10 * np.abs(u).max())


def _block_case(k, seed=0):
    nel = (16, 8, 8)
    bc, _ = problems.cantilever3d(*nel)
    x = np.random.default_rng(seed).uniform(0.1, 1.0, 16 * 64)
    K = fe.Assembly(*nel, fe.element_stiffness(3), bc=bc).matrix(fe.simp(x))
    B = np.random.default_rng(seed + 1).standard_normal((K.shape[0], k))
    B[bc, :] = 0.0
    return nel, K, B


def test_block_mgcg_matches_spsolve():
    """Block PCG (iterative.py:340-403 with a 2D rhs) converges every column to the direct solution
    and needs no more iterations than the hardest single column."""
    nel, K, B = _block_case(4)
    cg = CG(preconditioner=build_mg(nel, 3, omega=0.6, smooth_steps=2), tol=1e-12)
    cg.update(K)
    X = cg.solve(B)
    ref = spsolve(K.tocsc(), B)
    npt.assert_allclose(X, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())
    its_block = cg.iterations
    its_single = []
    for j in range(B.shape[1]):
        cg.solve(B[:, j])
        its_single.append(cg.iterations)
    assert its_block <= max(its_single)


def test_block_mgcg_rank_deficient_columns():
    """Linearly dependent columns are dropped by orth (iterative.py:283-285) without breakdown."""
    nel, K, B = _block_case(2, seed=3)
    B3 = np.stack([B[:, 0], 2.0 * B[:, 0], B[:, 1]], axis=1)
    cg = CG(preconditioner=build_mg(nel, 3, omega=0.6, smooth_steps=2), tol=1e-12)
    cg.update(K)
    X = cg.solve(B3)
    ref = spsolve(K.tocsc(), B3)
    npt.assert_allclose(X, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())


# ---------------------------------------------------------------- end-to-end vs the reference
def test_pipeline_c1_mbb_vs_reference():
    g = golden("c1_mbb.npz")
    bc, f = problems.mbb2d(120, 40)
    npt.assert_array_equal(bc, g["bc"])
    npt.assert_array_equal(f, g["f"])
    out = design_iteration((120, 40, 0), g["x"], bc, f, radius=2.0, plane="stress")
    npt.assert_allclose(out["xf"], g["xf"], rtol=1e-13)
    npt.assert_allclose(out["xsimp"], g["xsimp"], rtol=1e-13)
    npt.assert_allclose(out["u"], g["u"], rtol=1e-9, atol=1e-12 * np.abs(g["u"]).max())
    npt.assert_allclose(out["c"], float(g["c"]), rtol=1e-12)
    npt.assert_allclose(out["dcdx"], g["dcdx"], rtol=1e-9, atol=1e-12 * np.abs(g["dcdx"]).max())


@pytest.mark.parametrize("name", ["c2_direct_16x8x8.npz", "c2_direct_24x12x12.npz"])
def test_pipeline_c2_direct_vs_reference(name):
    g = golden(name)
    nel = tuple(int(v) for v in g["nel"])
    bc, f = problems.cantilever3d(*nel)
    npt.assert_array_equal(bc, g["bc"])
    npt.assert_array_equal(f, g["f"])
    out = design_iteration(nel, g["x"], bc, f, radius=1.5)
    npt.assert_allclose(out["xf"], g["xf"], rtol=1e-13)
    npt.assert_allclose(out["c"], float(g["c"]), rtol=1e-12)
    npt.assert_allclose(out["dcdx"], g["dcdx"], rtol=1e-9, atol=1e-12 * np.abs(g["dcdx"]).max())


@pytest.mark.parametrize("levels", [3, 4])
def test_pipeline_c2_mgcg_vs_reference(levels):
    g = golden(f"c2_mgcg_32x16x16_L{levels}.npz")
    nel = tuple(int(v) for v in g["nel"])
    bc, f = problems.cantilever3d(*nel)
    out = design_iteration(nel, g["x"], bc, f, radius=1.5, solver="mgcg", mg_levels=levels,
                           omega=float(g["omega"]), smooth_steps=int(g["steps"]), tol=float(g["tol"]))
    assert out["iterations"] == int(g["iterations"])
    # the golden history was parsed from the reference's printed residuals (about 6 significant digits)
    npt.assert_allclose(out["history"], g["history"], rtol=1e-3)
    assert out["adjoint_solved"] == bool(g["adjoint_solved"]) is False
    npt.assert_allclose(out["c"], float(g["c"]), rtol=1e-11)
    npt.assert_allclose(out["dcdx"], g["dcdx"], rtol=1e-8, atol=1e-11 * np.abs(g["dcdx"]).max())


def test_oc_step_volume_move_limits_and_uniform_case():
    from oracle.optim import oc_step
    rng = np.random.default_rng(0)
    x = rng.uniform(0.2, 0.8, 500)
    dg = -rng.uniform(0.1, 2.0, 500)
    xn = oc_step(x, dg, maxvol=0.5, move=0.2)
    assert np.all(xn >= np.maximum(0.0, x - 0.2) - 1e-15) and np.all(xn <= np.minimum(1.0, x + 0.2) + 1e-15)
    assert abs(np.mean(xn) - 0.5) < 1e-3
    # uniform sensitivity and design: the multiplier makes x*sqrt(-dg/l) = maxvol exactly
    xu = np.full(100, 0.4)
    xn = oc_step(xu, np.full(100, -1.0), maxvol=0.45, move=0.2)
    npt.assert_allclose(xn, 0.45, atol=1e-4)  # bisection tolerance 1e-4 on the multiplier


def test_density_filter_oracle_interior_matches_convolution_and_is_self_adjoint():
    """DensityFilter (filter.py:223-378): away from the boundary it is the normalised convolution of
    FilterConv (same weights max(0, r - d)); H is symmetric so <F x, y> = <x, H (y / s)>."""
    from oracle.filter import DensityFilterOracle
    nel = (12, 10, 8)
    rng = np.random.default_rng(4)
    x = rng.random(12 * 10 * 8)
    for r in (1.5, 2.0):
        F = DensityFilterOracle(nel, radius=r)
        y = F.forward(x)
        yc = ConvFilter(nel, filter_weights(r, nel)).forward(x)
        d = int(r)
        k, j, i = np.meshgrid(np.arange(8), np.arange(10), np.arange(12), indexing="ij")
        inner = ((i >= d) & (i < 12 - d) & (j >= d) & (j < 10 - d) & (k >= d) & (k < 8 - d)).ravel()
        npt.assert_allclose(y[inner], yc[inner], rtol=1e-13)
        assert abs(F.H - F.H.T).max() == 0.0
        dy = rng.random(x.size)
        npt.assert_allclose(np.dot(F.forward(x), dy), np.dot(x, F.adjoint(dy)), rtol=1e-12)


def test_heaviside_oracle_known_values_and_derivative():
    from oracle.pipeline import heaviside
    x = np.linspace(0.0, 1.0, 41)
    for eta in (0.3, 0.5, 0.7):
        y, dy = heaviside(x, 8.0, eta)
        npt.assert_allclose([y[0], y[-1]], [0.0, 1.0], atol=1e-15)  # end points are fixed for every beta, eta
        assert np.all(np.diff(y) > 0)
        e = 1e-6
        npt.assert_allclose(dy, (heaviside(x + e, 8.0, eta)[0] - heaviside(x - e, 8.0, eta)[0]) / (2 * e), rtol=1e-7)
    # eta = 0.5 is symmetric about (0.5, 0.5); beta -> 0 tends to the identity
    y, _ = heaviside(x, 8.0, 0.5)
    npt.assert_allclose(y + y[::-1], 1.0, atol=1e-14)
    npt.assert_allclose(heaviside(x, 1e-4, 0.3)[0], x, atol=1e-7)


def test_robust_iteration_oracle_vs_finite_differences():
    """c = sum_j f_j.u_j and its gradient through filter, Heaviside, SIMP and the multi-load solve,
    against central differences of c on a 4x2x2 grid (pins every factor of the adjoint chain)."""
    from oracle.pipeline import robust_iteration
    nel = (4, 2, 2)
    bc = problems.cantilever3d(*nel)[0]
    F = problems.face_point_loads(*nel, 3, 5)
    x = np.random.default_rng(2).uniform(0.2, 0.9, 16)
    out = robust_iteration(nel, x, bc, F, radius=1.5, etas=(0.3, 0.7), beta=4.0)
    for o in out:
        U = o["U"]
        K = fe.Assembly(*nel, fe.element_stiffness(3), bc).matrix(o["xsimp"])
        npt.assert_allclose(np.sum(U * (K @ U)), o["c"], rtol=1e-12)  # c = sum_j u_j^T K u_j
    e = 1e-6
    for i in (0, 5, 11, 15):
        xp, xm = x.copy(), x.copy()
        xp[i] += e
        xm[i] -= e
        cp = robust_iteration(nel, xp, bc, F, radius=1.5, etas=(0.3, 0.7), beta=4.0)
        cm = robust_iteration(nel, xm, bc, F, radius=1.5, etas=(0.3, 0.7), beta=4.0)
        for k, o in enumerate(out):
            fd = (cp[k]["c"] - cm[k]["c"]) / (2 * e)
            npt.assert_allclose(o["dcdx"][i], fd, rtol=2e-6)


# ---------------------------------------------------------------------------------------------
# scalar (Poisson) element and the partitioned solve with prescribed unknowns

@pytest.mark.parametrize("dim", [2, 3])
def test_poisson_element_one_element_heat_conduction(dim):
    """The reference's one-element wall (tests/test_assembly.py:286-340): T = 0 on the x = 0 face, heat
    Q spread over the x = Lx face; T there equals Q Lx / (k A) with A the cross-section (Ly Lz)."""
    Lx, Ly, Lz, k, Q = 2.0, 3.0, 4.0, 1.0, 1.0
    P0 = fe.element_poisson(dim, (Lx, Ly, Lz), k)
    nodes = np.arange(2 ** dim)
    left, right = nodes[(nodes & 1) == 0], nodes[(nodes & 1) == 1]
    asm = fe.Assembly(1, 1, 1 if dim == 3 else 0, P0, bc=left)
    q = np.zeros(2 ** dim)
    q[right] = Q / right.size
    T = np.linalg.solve(asm.matrix(np.ones(1)).toarray(), q)
    np.testing.assert_allclose(T[left], 0.0, atol=1e-12)
    np.testing.assert_allclose(T[right], Q * Lx / (k * Ly * Lz), rtol=1e-12)


def test_poisson_element_textbook_forms():
    # bilinear unit square: (1/6) [[4,-1,-1,-2], ...] in node order 0(-,-) 1(+,-) 2(-,+) 3(+,+), unit thickness
    sq = np.array([[4, -1, -1, -2], [-1, 4, -2, -1], [-1, -2, 4, -1], [-2, -1, -1, 4]]) / 6.0
    np.testing.assert_allclose(fe.element_poisson(2), sq, rtol=1e-14, atol=1e-15)
    # trilinear unit cube: 1/3 on the diagonal, 0 along an edge, -1/12 across a face and the body diagonal
    P = fe.element_poisson(3)
    for a in range(8):
        for b in range(8):
            hamming = bin(a ^ b).count("1")
            want = {0: 1 / 3, 1: 0.0, 2: -1 / 12, 3: -1 / 12}[hamming]
            assert abs(P[a, b] - want) < 1e-14
    # scaling: k and the element size (2D: thickness h_z multiplies, square shape invariant in 2D)
    np.testing.assert_allclose(fe.element_poisson(2, (2.0, 2.0, 3.0), 5.0), 15.0 * sq, rtol=1e-14)
    np.testing.assert_allclose(fe.element_poisson(3, (2.0, 2.0, 2.0), 1.0), 2.0 * P, rtol=1e-14)
    for d in (2, 3):
        Pd = fe.element_poisson(d, (0.7, 1.3, 2.1), 1.7)
        np.testing.assert_allclose(Pd, Pd.T, atol=1e-15)
        np.testing.assert_allclose(Pd @ np.ones(2 ** d), 0.0, atol=1e-14)  # constants in the null space


def _fd_check(fun, x, grad, h=1e-6, rtol=1e-5, n=6, seed=0):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        dx = rng.standard_normal(x.shape)
        fd = (fun(x + h * dx) - fun(x - h * dx)) / (2 * h)
        an = float(np.sum(grad * dx))
        assert abs(fd - an) <= rtol * max(abs(fd), abs(an), 1e-12), (fd, an)


@pytest.mark.parametrize("nel,physics", [((10, 10, 0), "elastic"), ((6, 4, 4), "elastic"), ((8, 6, 0), "thermal")])
def test_system_of_equations_oracle(nel, physics):
    """SystemOfEquations (linalg.py:222-349): identities and finite differences of g = c1 . x + c2 . b through
    the densities, the applied load and the prescribed values (the reference tests test_linsolve_sparse.py:206-420
    check the same three gradients by finite differences)."""
    from oracle.pipeline import SystemOfEquations
    nelx, nely, nelz = nel
    dim = fe.grid_dim(*nel)
    K0 = fe.element_poisson(dim) if physics == "thermal" else fe.element_stiffness(dim)
    ndof = K0.shape[0] // 2 ** dim
    asm = fe.Assembly(nelx, nely, nelz, K0)
    n = asm.n
    nny, nnz = nely + 1, max(nelz, 0) + 1
    left = np.array([(k * nny + j) * (nelx + 1) for k in range(nnz) for j in range(nny)])
    right = left + nelx
    # every unknown of the x = 0 face and the last one of the x = max face (no rigid-body mode left free)
    p = np.unique(np.concatenate([(left[:, None] * ndof + np.arange(ndof)).ravel(), right * ndof + ndof - 1]))
    f = np.setdiff1d(np.arange(n), p)
    rng = np.random.default_rng(1)
    x0 = rng.uniform(0.1, 1.0, asm.dc.shape[0])
    bf0 = rng.standard_normal(f.size)
    xp0 = rng.standard_normal(p.size)
    c1, c2 = rng.standard_normal(n), rng.standard_normal(n)

    def run(xe, bf, xp):
        soe = SystemOfEquations(free=f, prescribed=p)
        X, B = soe.forward(asm.matrix(xe), bf, xp)
        return soe, X, B

    soe, X, B = run(x0, bf0, xp0)
    A = asm.matrix(x0)
    np.testing.assert_allclose(A @ X, B, rtol=1e-9, atol=1e-9 * np.abs(B).max())
    np.testing.assert_array_equal(X[p], xp0)
    np.testing.assert_array_equal(B[f], bf0)
    lam, xx, dbf, dxp = soe.sensitivity(c1, c2)
    dxe = asm.sens_contract([lam], [xx])

    def g(xe=x0, bf=bf0, xp=xp0):
        _, X_, B_ = run(xe, bf, xp)
        return float(c1 @ X_ + c2 @ B_)
    _fd_check(lambda v: g(xe=v), x0, dxe)
    _fd_check(lambda v: g(bf=v), bf0, dbf, rtol=1e-6)
    _fd_check(lambda v: g(xp=v), xp0, dxp, rtol=1e-6)
