"""Host-side checks that need no GPU: C-ABI exports, numbering, setup-time maps, problem data."""
import os
import re

import numpy as np
import numpy.testing as npt
import pytest

import pymoto_amd as pm
from conftest import ROOT, golden
from oracle import fe, problems
from oracle.filter import ConvFilter, filter_weights


def test_library_exports_every_declared_symbol():
    from pymoto_amd import _lib
    lib = _lib.load()
    header = open(os.path.join(ROOT, "include", "pymoto_amd.h")).read()
    declared = set(re.findall(r"\b(pmt_[a-z0-9_]+)\s*\(", header))
    assert declared, "no declarations found"
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared in include/pymoto_amd.h but not exported"
    assert declared == set(_lib.exported_symbols())
    assert lib.pmt_abi_version() == 1


def test_host_helpers_of_the_library():
    from pymoto_amd import _lib
    lib = _lib.load()
    assert lib.pmt_csr_nnz(120, 40, 0, 2) == 174724
    g = golden("c2_direct_16x8x8.npz")
    assert lib.pmt_csr_nnz(16, 8, 8, 3) == int(g["nnz"])
    assert lib.pmt_reduce_work_size() >= 1024


def _k0_extend(K0):
    from pymoto_amd import _lib
    lib = _lib.load()
    dim = 3 if K0.shape[0] == 24 else 2
    K0 = np.ascontiguousarray(K0, dtype=np.float64)
    ext = np.zeros(_lib.K0_EXT_LEN)
    bits = lib.pmt_k0_extend(dim, K0.ctypes.data_as(_lib.c_vp), ext.ctypes.data_as(_lib.c_vp))
    return bits, ext


def _reflection_product(tab, u):
    """K0 u from the 8 reflection-basis blocks: WHT over the corner bits, 3x3 block per character
    chi (dof (d, k) belongs to chi = k ^ 2^d), WHT back."""
    H = np.array([[(-1.0) ** bin(k & a).count("1") for a in range(8)] for k in range(8)])
    U = H @ u.reshape(8, 3)  # U[k, d]
    W = np.zeros_like(U)
    for chi in range(8):
        vin = np.array([U[chi ^ (1 << d), d] for d in range(3)])
        out = tab[chi * 9:(chi + 1) * 9].reshape(3, 3) @ vin
        for d in range(3):
            W[chi ^ (1 << d), d] = out[d]
    return (H @ W).reshape(-1)


@pytest.mark.parametrize("esize,nu", [((1.0, 1.0, 1.0), 0.3), ((1.0, 2.0, 0.5), 0.3), ((1.0, 1.0, 1.0), 0.45)])
def test_k0_reflection_form(esize, nu):
    """pmt_k0_extend: the block tables reproduce K0 u (identity of the mid-plane reflection basis) for
    box elements; the FP32 copy is the rounded FP64 table; K0 itself leads the buffer."""
    K0 = fe.element_stiffness(3, esize=esize, nu=nu)
    bits, ext = _k0_extend(K0)
    # box elements: class pattern, reflection form, rigid-translation entries zero; the cube also the
    # 8-number isotropic form (PMT_K0_CUBE)
    assert bits == (15 if esize == (1.0, 1.0, 1.0) else 7)
    npt.assert_array_equal(ext[:576], K0.ravel())
    tab = ext[576:648]
    npt.assert_array_equal(ext[648:684].view(np.float32), tab.astype(np.float32))
    rng = np.random.default_rng(5)
    for _ in range(4):
        u = rng.standard_normal(24)
        ref = K0 @ u
        npt.assert_allclose(_reflection_product(tab, u), ref, rtol=0, atol=1e-14 * np.abs(K0).max() * np.abs(u).sum())
    # rigid translations are in the null space: their blocks have an exact zero row
    for d in range(3):
        t = np.zeros((8, 3))
        t[:, d] = 1.0
        assert np.abs(_reflection_product(tab, t.ravel())).max() < 1e-14 * np.abs(K0).max()


def test_k0_without_reflection_symmetry_is_refused_the_form():
    K0 = fe.element_stiffness(3)
    rng = np.random.default_rng(1)
    P = rng.standard_normal((24, 24)) * 1e-6
    bits, ext = _k0_extend(K0 + P + P.T)
    assert bits & 2 == 0
    assert not ext[576:].any()
    bits2, ext2 = _k0_extend(fe.element_stiffness(2))
    assert bits2 == 1 and not ext2[64:].any()


def test_errors_are_reported_without_gpu():
    from pymoto_amd import _lib
    lib = _lib.load()
    rc = lib.pmt_restrict(3, 2, 2, None, None, None)
    assert rc == -1
    assert "even" in lib.pmt_last_error().decode()


def test_domain_numbering_matches_oracle():
    for nel in [(5, 4, 0), (3, 4, 5)]:
        d = pm.VoxelDomain(*nel)
        npt.assert_array_equal(d.conn, fe.elem_connectivity(*nel))
        npt.assert_array_equal(d.get_dofconnectivity(d.dim), fe.dofconn(*nel))
    d = pm.VoxelDomain(100, 142, 284, unitx=0.1, unity=0.2, unitz=0.3)
    n = d.get_nodenumber(10, 20, 30)
    assert tuple(d.get_node_indices(n)) == (10, 20, 30)
    npt.assert_allclose(d.get_node_position(n), [1.0, 4.0, 9.0])


def test_deprecated_alias():
    with pytest.warns(DeprecationWarning):
        assert pm.DomainDefinition is pm.VoxelDomain


def test_element_stiffness_matches_golden():
    from pymoto_amd.assembly import element_stiffness
    g = golden("k0_cases.npz")
    for i in range(int(g["ncases"])):
        is3d, sx, sy, sz = g[f"spec_{i}"]
        dom = pm.VoxelDomain(1, 1, 1 if is3d else 0, unitx=sx, unity=sy, unitz=sz)
        K0 = element_stiffness(dom, 2.5, 0.27, str(g[f"plane_{i}"]))
        npt.assert_allclose(K0, g[f"K0_{i}"], rtol=1e-13, atol=1e-13 * np.abs(K0).max())


def test_problems_match_oracle():
    from pymoto_amd import problems as pp
    for nel in [(120, 40, 0), (16, 8, 8)]:
        d = pm.VoxelDomain(*nel)
        if d.dim == 2:
            bc, f = pp.mbb2d(d)
            obc, of = problems.mbb2d(nel[0], nel[1])
        else:
            bc, f = pp.cantilever3d(d)
            obc, of = problems.cantilever3d(*nel)
        npt.assert_array_equal(bc, obc)
        npt.assert_array_equal(f, of)


def test_bc_mask():
    from pymoto_amd.assembly import bc_node_mask
    d = pm.VoxelDomain(3, 3)
    m = bc_node_mask(d, 2, [1, 2, 8, 15, 2])
    assert m[0] == 2 and m[1] == 1 and m[4] == 1 and m[7] == 2 and m.sum() == 6


@pytest.mark.parametrize("case", range(8))
def test_filter_axis_maps_reproduce_reference_padding(case):
    """The per-axis maps must rebuild exactly the reference's padded element-index array."""
    from pymoto_amd.filter import FilterConv
    g = golden("filter_cases.npz")
    nel = tuple(int(v) for v in g[f"case{case}_nel"])
    raw = [str(s) for s in g[f"case{case}_bcs"]]
    bcs = [s if s in ("symmetric", "edge", "wrap") else float(s) for s in raw]
    w = g[f"case{case}_weights"]
    m = FilterConv(pm.VoxelDomain(*nel), weights=w, xmin_bc=bcs[0], xmax_bc=bcs[1], ymin_bc=bcs[2],
                   ymax_bc=bcs[3], zmin_bc=bcs[4], zmax_bc=bcs[5])
    ref = ConvFilter(nel, w, ((bcs[0], bcs[1]), (bcs[2], bcs[3]), (bcs[4], bcs[5])))
    sizes = [nel[0], nel[1], max(nel[2], 1)]
    L = [sizes[a] + w.shape[a] - 1 for a in range(3)]
    mx, my, mz = np.split(m._map, [L[0], L[0] + L[1]])
    cx, cy, cz = np.split(m._cval, [L[0], L[0] + L[1]])
    # rebuild the padded field for a random x and compare with the reference's padded vector
    x = np.random.default_rng(case).random(int(np.prod(sizes)))
    xpad_ref = x[ref.el3d_pad]
    for ix, val in ref.overrides:
        xpad_ref[ix] = val
    A, B, C = np.meshgrid(np.arange(L[0]), np.arange(L[1]), np.arange(L[2]), indexing="ij")
    ma, mb, mc = mx[A], my[B], mz[C]
    xv = x[((np.maximum(mc, 0) * sizes[1] + np.maximum(mb, 0)) * sizes[0] + np.maximum(ma, 0))]
    xv = np.where(ma < 0, cx[A], xv)
    xv = np.where(mb < 0, cy[B], xv)
    xv = np.where(mc < 0, cz[C], xv)
    npt.assert_array_equal(xv, xpad_ref)
    if np.isfinite(float(g[f"case{case}_radius"])):
        npt.assert_allclose(filter_weights(float(g[f"case{case}_radius"]), nel), w, rtol=1e-15)


def test_powlaw_parse_matches_lambdify_constants():
    from pymoto_amd.generic import _parse_powlaw
    a, b, p, c, q = _parse_powlaw(f"{1e-9} + {1.0 - 1e-9}*inp0^3")
    assert (a, b, p, c, q) == (1e-9, 0.999999999, 3.0, 2.999999997, 2.0)
    with pytest.raises(NotImplementedError):
        _parse_powlaw("sin(inp0)")


def test_auto_levels():
    from pymoto_amd.solvers import auto_levels
    # the reference example's rule (ex_compliance_multigrid.py:108-121): coarse extents even and >= 8
    assert auto_levels(pm.VoxelDomain(120, 40)) == 4
    assert auto_levels(pm.VoxelDomain(256, 128, 128)) == 6
    assert auto_levels(pm.VoxelDomain(192, 96, 96)) == 5
    assert auto_levels(pm.VoxelDomain(16, 8, 8)) == 2


def test_device_modules_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    d = pm.VoxelDomain(4, 2)
    with pytest.raises(Exception):
        pm.FilterConv(d, radius=1.5)(pm.Signal("x", np.ones(d.nel)))


def _kernel_indptr(nelx, nely, nelz, D):
    """Literal Python transcription of csr_pattern_kernel's closed-form row pointer (csrc/csr.hip)."""
    nx, ny, nz = nelx + 1, nely + 1, (nelz + 1 if nelz > 0 else 1)

    def cnt(i, n):
        return 1 if n == 1 else 3 - (i == 0) - (i == n - 1)

    def pre(i, n):
        if i <= 0:
            return 0
        if n == 1:
            return i
        return 3 * i - 1 - (1 if i > n - 1 else 0)

    Sx, Sy, Sz = pre(nx, nx), pre(ny, ny), pre(nz, nz)
    out = []
    for n in range(nx * ny * nz):
        i, j, k = n % nx, (n // nx) % ny, n // (nx * ny)
        before = pre(k, nz) * Sy * Sx + cnt(k, nz) * pre(j, ny) * Sx + cnt(k, nz) * cnt(j, ny) * pre(i, nx)
        for d in range(D):
            out.append(D * D * before + d * D * cnt(i, nx) * cnt(j, ny) * cnt(k, nz))
    out.append(D * D * Sx * Sy * Sz)
    return np.array(out)


@pytest.mark.parametrize("nel", [(6, 4, 0), (1, 1, 0), (3, 1, 0), (1, 5, 0), (4, 3, 2), (1, 1, 1), (2, 1, 3), (5, 4, 3)])
def test_csr_closed_form_indptr_matches_oracle(nel):
    dim = 3 if nel[2] > 0 else 2
    asm = fe.Assembly(*nel, fe.element_stiffness(dim))
    npt.assert_array_equal(_kernel_indptr(*nel, dim), asm.indptr)
    from pymoto_amd import _lib
    assert _lib.load().pmt_csr_nnz(*nel, dim) == asm.indices.size


def test_heaviside_expression_parsing():
    """The robust formulation's projection (ex_compliance_robust.py:139) is recognised with beta as a
    second input or a constant; SIMP is not mistaken for it."""
    from pymoto_amd.generic import _parse_heaviside
    h = "(tanh(inp1 * {0}) + tanh(inp1 * (inp0 - {0}))) / (tanh(inp1 * {0}) + tanh(inp1 * (1 - {0})))"
    assert _parse_heaviside(h.format(0.3)) == (None, 0.3)
    assert _parse_heaviside(h.format(0.7).replace("inp1", "8")) == (8.0, 0.7)
    assert _parse_heaviside(f"{1e-9} + {1 - 1e-9}*inp0^3") is None
    assert _parse_heaviside("tanh(inp0)") is None


def test_vti_writer_matches_oracle_bytes_and_decodes(tmp_path):
    """write_to_vti (domain.py:452-604): same bytes as the oracle restatement for 2D/3D point and cell
    data (2-component padding, (n, k) blocks), and the arrays decode back to the inputs."""
    from oracle.io import read_vti_arrays, write_vti
    rng = np.random.default_rng(0)
    for nel in [(7, 5, 0), (4, 3, 2)]:
        d = pm.VoxelDomain(*nel)
        vecs = {"x": rng.random(d.nel), "u": rng.standard_normal(d.dim * d.nnodes),
                "U": rng.standard_normal((d.dim * d.nnodes, 3)), "S": rng.random((2, d.nel))}
        a, b = tmp_path / f"a{nel[2]}.vti", tmp_path / f"b{nel[2]}.vti"
        import pymoto_amd.io as pio
        chunk = pio._CHUNK
        pio._CHUNK = 3 * 8  # stream in many small chunks
        try:
            d.write_to_vti(vecs, str(a), scale=2.0)
        finally:
            pio._CHUNK = chunk
        write_vti(d, vecs, str(b), scale=2.0)
        assert a.read_bytes() == b.read_bytes()
        arrays = read_vti_arrays(str(a))
        npt.assert_allclose(arrays["x"][1], vecs["x"].astype(np.float32))
        u = arrays["u"][1].reshape(-1, 3)
        npt.assert_allclose(u[:, : d.dim].ravel(), vecs["u"].astype(np.float32))
        npt.assert_allclose(arrays["U(1)"][1].reshape(-1, 3)[:, : d.dim].ravel(), vecs["U"][:, 1].astype(np.float32))
        npt.assert_allclose(arrays["S(1)"][1], vecs["S"][1].astype(np.float32))


def test_face_point_loads_match_oracle():
    from pymoto_amd import problems as pp
    d = pm.VoxelDomain(12, 6, 4)
    npt.assert_array_equal(pp.face_point_loads(d, 4, 7), problems.face_point_loads(12, 6, 4, 4, 7))
    F = pp.face_point_loads(d, 4, 7)
    assert np.all(np.sum(F, axis=0) == -1.0) and np.count_nonzero(F) == 4


def _vti_golden_cases():
    from conftest import golden
    g = golden("vti_cases.npz")
    for i in range(int(g["ncases"])):
        nel = tuple(int(v) for v in g[f"nel_{i}"])
        unit = g[f"unit_{i}"]
        vecs = {k: g[f"{k}_{i}"] for k in ("x", "u", "modes")}
        yield nel, unit, float(g[f"scale_{i}"]), vecs, g[f"bytes_{i}"].tobytes()


def test_vti_writer_bytes_equal_reference_golden(tmp_path):
    """The bytes of the reference's own VoxelDomain.write_to_vti (domain.py:452-604; fixture written by
    tests/golden/make_golden.py) for a 2D and a 3D grid: cell scalar, nodal vector (2D padded to 3
    components) and a block of 3 nodal vectors - from this package's writer (numpy inputs) and from the
    oracle's restatement."""
    from oracle.io import write_vti
    for nel, unit, scale, vecs, want in _vti_golden_cases():
        d = pm.VoxelDomain(*nel, unitx=unit[0], unity=unit[1], unitz=unit[2])
        a, b = tmp_path / f"pkg{nel[2]}.vti", tmp_path / f"orc{nel[2]}.vti"
        d.write_to_vti(vecs, str(a), scale=scale)
        write_vti(d, vecs, str(b), scale=scale)
        assert a.read_bytes() == want
        assert b.read_bytes() == want
