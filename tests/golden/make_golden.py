"""Write golden input/output vectors of the reference pyMOTO v2.0.1 for the FE hot path.

Runs ONLY in a container where the reference is mounted at /root/reference; it imports it from
there and writes small ``.npz`` fixtures next to this file.  Nothing else in the repository
imports the reference; the fixtures are data (inputs and the reference's outputs).

Why an import is needed for these (no closed form or library routine pins them exactly):
* end-to-end values of the reference's FP64 path on seeded inputs (sympy-lambdified SIMP
  constants, LDAS adjoint reuse, SuperLU): ``xf``, ``xsimp``, ``u``, ``c``, ``dc/dx``;
* the exact CSR ``indptr``/``indices`` (and their integer dtype) and the bit pattern of the
  ``np.add.at``-assembled CSR data;
* the iteration count and residual history of the reference's own GeometricMultigrid + CG;
* FilterConv outputs for the 'edge' / 'wrap' / constant paddings and non-symmetric kernels,
  which the reference's own tests do not pin;
* the bytes of the VTI files written by VoxelDomain.write_to_vti (a file format: no formula pins it).

Usage:  MPLBACKEND=Agg python tests/golden/make_golden.py [case ...]
"""
import hashlib
import os
import sys

import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(OUT), ".."))  # repo root for oracle.problems
sys.path.insert(0, REF)

import scipy.sparse as sps  # noqa: E402
import pymoto as pym  # noqa: E402
from oracle.problems import cantilever3d, mbb2d  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def compliance_chain(domain, x, bc, f, radius, xmin=1e-9, plane="strain", solver=None, filter_kw=None):
    sx = pym.Signal("x", state=x.copy())
    with pym.Network() as fn:
        sxf = pym.FilterConv(domain, radius=radius, **(filter_kw or {}))(sx)
        ssimp = pym.MathExpression(f"{xmin} + {1.0 - xmin}*inp0^3")(sxf)
        sK = pym.AssembleStiffness(domain, bc=bc, plane=plane)(ssimp)
        lin = pym.LinSolve(hermitian=True, solver=solver)
        su = lin(sK, f)
        sc = pym.EinSum("i,i->")(su, f)
    fn.response()
    sc.sensitivity = 1.0
    fn.sensitivity()
    return dict(x=x, xf=sxf.state, xsimp=ssimp.state, u=su.state, c=np.float64(sc.state),
                dcdx=sx.sensitivity, dcdxsimp=ssimp.sensitivity, K=sK.state, lin=lin)


def c1_mbb():
    nelx, nely = 120, 40
    dom = pym.VoxelDomain(nelx, nely)
    x = np.random.default_rng(1).uniform(0.01, 1.0, dom.nel)
    bc, f = mbb2d(nelx, nely)
    r = compliance_chain(dom, x, bc, f, radius=2.0, plane="stress")
    K = r["K"]
    K0 = pym.AssembleStiffness(dom, plane="stress").stiffness_element
    np.savez_compressed(os.path.join(OUT, "c1_mbb.npz"), nel=np.array([nelx, nely, 0]), seed=1,
                        x=x, xf=r["xf"], xsimp=r["xsimp"], u=r["u"], c=r["c"], dcdx=r["dcdx"],
                        dcdxsimp=r["dcdxsimp"], bc=bc, f=f, K0=K0,
                        indptr=K.indptr, indices=K.indices, data=K.data,
                        index_dtype=str(K.indices.dtype),
                        sha_indptr=sha(K.indptr.astype(np.int64)), sha_indices=sha(K.indices.astype(np.int64)),
                        sha_data=sha(K.data))


def c2_small_direct(nel, seed):
    dom = pym.VoxelDomain(*nel)
    x = np.random.default_rng(seed).uniform(0.0, 1.0, dom.nel)
    bc, f = cantilever3d(*nel)
    r = compliance_chain(dom, x, bc, f, radius=1.5)
    K = r["K"]
    name = "c2_direct_{}x{}x{}.npz".format(*nel)
    np.savez_compressed(os.path.join(OUT, name), nel=np.array(nel), seed=seed, x=x, xf=r["xf"],
                        xsimp=r["xsimp"], u=r["u"], c=r["c"], dcdx=r["dcdx"], dcdxsimp=r["dcdxsimp"],
                        bc=bc, f=f, K0=pym.AssembleStiffness(dom).stiffness_element,
                        index_dtype=str(K.indices.dtype), nnz=K.nnz,
                        sha_indptr=sha(K.indptr.astype(np.int64)), sha_indices=sha(K.indices.astype(np.int64)),
                        sha_data=sha(K.data))


def c2_small_mgcg(nel, seed, levels=3, omega=0.6, steps=2, tol=1e-8):
    dom = pym.VoxelDomain(*nel)
    x = np.random.default_rng(seed).uniform(0.0, 1.0, dom.nel)
    bc, f = cantilever3d(*nel)
    mgs = [pym.solvers.GeometricMultigrid(dom, smoother=pym.solvers.DampedJacobi(w=omega), smooth_steps=steps)]
    for _ in range(levels - 2):
        mgs.append(pym.solvers.GeometricMultigrid(mgs[-1].sub_domain, smoother=pym.solvers.DampedJacobi(w=omega),
                                                  smooth_steps=steps))
        mgs[-2].inner_level = mgs[-1]
    cg = pym.solvers.CG(preconditioner=mgs[0], tol=tol, verbosity=2)
    import io
    import contextlib
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        r = compliance_chain(dom, x, bc, f, radius=1.5, solver=cg)
    # first CG call only (the forward solve); later calls would be LDAS-triggered adjoint solves
    lines = buf.getvalue().splitlines()
    starts = [i for i, line in enumerate(lines) if "Initial" in line]
    first = lines[starts[0]:(starts[1] if len(starts) > 1 else len(lines))]
    hist = [float(line.split("=")[-1].strip().strip("[]").split()[0])
            for line in first if line.startswith("CG i =") or "Initial" in line]
    iters = sum(1 for line in first if line.startswith("CG i ="))
    adjoint_solved = bool(np.any(r["lin"].solver._did_solve))
    name = "c2_mgcg_{}x{}x{}_L{}.npz".format(*nel, levels)
    np.savez_compressed(os.path.join(OUT, name), nel=np.array(nel), seed=seed, levels=levels, omega=omega,
                        steps=steps, tol=tol, x=x, xf=r["xf"], u=r["u"], c=r["c"], dcdx=r["dcdx"],
                        bc=bc, f=f, iterations=iters, history=np.array(hist), n_cg_calls=len(starts),
                        adjoint_solved=adjoint_solved)


def filter_cases():
    cases = {}
    rng = np.random.default_rng(7)
    specs = [
        ((10, 12, 0), dict(radius=2.0), {}),
        ((10, 12, 0), dict(radius=2.0), dict(xmin_bc="edge", xmax_bc="wrap", ymin_bc=0.0, ymax_bc=1.0)),
        ((9, 7, 0), dict(radius=3.2), dict(xmin_bc="wrap", xmax_bc="wrap", ymin_bc="edge", ymax_bc=0.5)),
        ((10, 21, 0), dict(weights=1.0 + np.arange(3).reshape((3, 1, 1))), {}),
        ((8, 7, 6), dict(radius=1.5), {}),
        ((8, 7, 6), dict(radius=2.3), dict(xmin_bc=0.0, zmax_bc="edge", ymin_bc="wrap", ymax_bc="wrap")),
        ((6, 6, 6), dict(radius=5.0), {}),
        ((5, 6, 7), dict(weights=rng.random((3, 5, 3))), dict(zmin_bc=0.25, xmax_bc="edge")),
    ]
    for i, (nel, kw, bcs) in enumerate(specs):
        dom = pym.VoxelDomain(*nel)
        x = rng.random(dom.nel)
        dy = rng.random(dom.nel)
        m = pym.FilterConv(dom, **kw, **bcs)
        sx = pym.Signal("x", state=x)
        sy = m(sx)
        sy.sensitivity = dy
        m.sensitivity()
        cases[f"case{i}_nel"] = np.array(nel)
        cases[f"case{i}_weights"] = m.weights
        cases[f"case{i}_bcs"] = np.array([str(bcs.get(k, "symmetric")) for k in
                                          ("xmin_bc", "xmax_bc", "ymin_bc", "ymax_bc", "zmin_bc", "zmax_bc")])
        cases[f"case{i}_radius"] = np.float64(kw.get("radius", np.nan))
        cases[f"case{i}_x"] = x
        cases[f"case{i}_y"] = sy.state
        cases[f"case{i}_dy"] = dy
        cases[f"case{i}_dx"] = sx.sensitivity
    cases["ncases"] = len(specs)
    np.savez_compressed(os.path.join(OUT, "filter_cases.npz"), **cases)


def k0_cases():
    out = {}
    specs = [((1, 1, 0), (1.0, 1.0, 1.0), "strain"), ((1, 1, 0), (1.0, 1.0, 1.0), "stress"),
             ((1, 1, 0), (0.1, 0.2, 0.3), "stress"), ((1, 1, 1), (1.0, 1.0, 1.0), "strain"),
             ((1, 1, 1), (0.1, 0.2, 0.3), "strain")]
    for i, (nel, size, plane) in enumerate(specs):
        dom = pym.VoxelDomain(*nel, unitx=size[0], unity=size[1], unitz=size[2])
        out[f"K0_{i}"] = pym.AssembleStiffness(dom, plane=plane, e_modulus=2.5, poisson_ratio=0.27).stiffness_element
        out[f"spec_{i}"] = np.array([nel[2] > 0, *size], dtype=float)
        out[f"plane_{i}"] = np.array(plane)
    out["ncases"] = len(specs)
    np.savez_compressed(os.path.join(OUT, "k0_cases.npz"), **out)


def small_csr_cases():
    """Full CSR arrays (pattern + bit-exact data) on small grids with BCs, for the device assembly."""
    out = {}
    specs = [((6, 4, 0), "stress", 3), ((4, 3, 2), "strain", 4), ((5, 4, 3), "strain", 5)]
    for i, (nel, plane, seed) in enumerate(specs):
        dom = pym.VoxelDomain(*nel)
        x = np.random.default_rng(seed).uniform(0.0, 1.0, dom.nel)
        if nel[2] == 0:
            bc, _ = mbb2d(nel[0], nel[1])
        else:
            bc, _ = cantilever3d(*nel)
        K = pym.AssembleStiffness(dom, bc=bc, plane=plane)(x)
        out[f"nel_{i}"] = np.array(nel)
        out[f"plane_{i}"] = np.array(plane)
        out[f"x_{i}"] = x
        out[f"bc_{i}"] = bc
        out[f"indptr_{i}"] = K.indptr
        out[f"indices_{i}"] = K.indices
        out[f"data_{i}"] = K.data
    out["ncases"] = len(specs)
    np.savez_compressed(os.path.join(OUT, "csr_cases.npz"), **out)


def vti_cases():
    """Bytes of VoxelDomain.write_to_vti (domain.py:452-604) for a 2D and a 3D grid: a cell scalar, a
    nodal vector (2D: padded to 3 components), a (k, ndof) block of nodal vectors and a scaled grid.
    The file format is pinned by nothing else (no closed form), so its bytes are the fixture."""
    import tempfile
    out = {}
    specs = [((5, 3, 0), 1.0, 11), ((4, 3, 2), 0.5, 12)]
    for i, (nel, scale, seed) in enumerate(specs):
        dom = pym.VoxelDomain(*nel, unitx=1.0, unity=0.5, unitz=2.0)
        rng = np.random.default_rng(seed)
        vecs = {"x": rng.uniform(0.0, 1.0, dom.nel),
                "u": rng.standard_normal(dom.nnodes * dom.dim),
                "modes": rng.standard_normal((3, dom.nnodes * dom.dim))}
        with tempfile.TemporaryDirectory() as tmp:
            fn = os.path.join(tmp, "case.vti")
            dom.write_to_vti(vecs, filename=fn, scale=scale)
            with open(fn, "rb") as fh:
                data = fh.read()
        out[f"nel_{i}"] = np.array(nel)
        out[f"unit_{i}"] = np.array([1.0, 0.5, 2.0])
        out[f"scale_{i}"] = np.array(scale)
        for k, v in vecs.items():
            out[f"{k}_{i}"] = v
        out[f"bytes_{i}"] = np.frombuffer(data, dtype=np.uint8)
    out["ncases"] = len(specs)
    np.savez_compressed(os.path.join(OUT, "vti_cases.npz"), **out)


if __name__ == "__main__":
    if len(sys.argv) > 1:  # only the named cases, e.g. `make_golden.py vti_cases`
        for name in sys.argv[1:]:
            globals()[name]()
        sys.exit(0)
    vti_cases()
    c1_mbb()
    c2_small_direct((16, 8, 8), seed=2)
    c2_small_direct((24, 12, 12), seed=3)
    c2_small_mgcg((32, 16, 16), seed=4, levels=3)
    c2_small_mgcg((32, 16, 16), seed=4, levels=4)
    filter_cases()
    k0_cases()
    small_csr_cases()
    print("golden fixtures written to", OUT)
