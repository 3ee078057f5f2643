"""Design-iteration benchmark: 3D cantilever, 256x128x128 hex8 (12,830,211 dofs), MG-PCG to 1e-8.

One step = one topology-optimisation design iteration through the pyMOTO-API device modules:
FilterConv(r=1.5) -> SIMP (xmin=1e-9, p=3) -> AssembleStiffness -> LinSolve (LDAS + MG-PCG,
tol 1e-8, omega 0.6, 2+2 Jacobi sweeps, warm start from the previous solution) -> compliance
EinSum -> full adjoint back to dc/dx.  Every step gets a fresh seeded density field (uploaded
before the timed region).  Solver variants, timed one after the other on the same designs:

* ``fp64``: the reference's arithmetic throughout (FP64 V-cycle and CG);
* ``fp32``: FP32 V-cycle (operator, smoother, transfers, stencils, coarsest inverse) inside the
  FP64 CG, whose residual still reaches 1e-8; compliance and dc/dx are compared with ``fp64``;
* ``bf16``: as ``fp32``, with the fine-grid operator of the V-cycle on the matrix cores (BF16
  element products, FP32 accumulation).

A variant may name its own Jacobi damping as ``precision@omega``.  ``fp64`` is the reference
configuration (omega = 0.6); the other default variants use omega = 0.5: with these SIMP fields
lambda_max(D^-1 K) is 3.5-3.8, so omega = 0.6 lies above the Jacobi stability bound 2/lambda_max and
the smoother amplifies the highest modes - the FP64 V-cycle still preconditions CG, but needs about
four times the iterations, and reduced-precision V-cycles degrade further.  ``speedup_vs_fp64`` is
against FP64 with the same omega as the fastest variant (the precision gain alone),
``speedup_vs_fp64_reference_omega`` against the reference configuration.

``value`` is the seconds per design iteration of the fastest variant whose compliance agrees with
the FP64 variant to 1e-7; every variant's time, CG iterations and errors are in ``config``.
Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nel 256 128 128] [--levels L] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--nel", type=int, nargs=3, default=[256, 128, 128])
    p.add_argument("--levels", type=int, default=0, help="grids in the MG hierarchy (0 = reference example rule)")
    p.add_argument("--tol", type=float, default=1e-8)
    p.add_argument("--omega", type=float, default=0.6, help="Jacobi damping of variants without '@omega'")
    p.add_argument("--variants", default="fp64,fp64@0.5,fp32@0.5,bf16@0.5,fp32@0.5+learned",
                   help="comma-separated precision[@omega][+learned] list; the first fp64 entry is the reference")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--no-batched", action="store_true", help="skip the batched multi-load workload")
    p.add_argument("--learned-samples", type=int, default=24, help="training pairs for '+learned' variants")
    p.add_argument("--learned-epochs", type=int, default=4)
    p.add_argument("--c5-pairs", type=int, default=2000,
                   help="solver-generated OC-trajectory pairs of the 128x64x64 learned workload (0 = skip it)")
    p.add_argument("--c5-epochs", type=int, default=2)
    p.add_argument("--cpu-nel", type=int, nargs=3, default=[96, 48, 48])
    p.add_argument("--ku-only", action="store_true",
                   help="kernel-only driver for hardware counters: 5 fine FP64 K.u launches, then 5 copies of one "
                        "nodal vector (calibration of the byte counters); prints no JSON line")
    p.add_argument("--kernels-only", action="store_true",
                   help="kernel-only driver: 5 launches each of the fine K.u (FP64, FP32, BF16), V-cycles (FP64, "
                        "FP32), FilterConv forward and adjoint, SIMP forward and adjoint, the sensitivity contraction "
                        "and the learned network's forward pass, then 5 copies of one nodal vector; prints no JSON "
                        "line")
    return p.parse_args()


def parse_variant(v):
    """'fp32', 'fp32@0.5', 'fp32@cheb' or 'fp32@0.5+learned' -> (precision, omega, "chebyshev" or None)."""
    prec, _, om = v.split("+")[0].partition("@")
    if om == "cheb":
        return prec.strip(), "chebyshev"
    return prec.strip(), (float(om) if om else None)


def smoother_name(omega):
    return "Chebyshev smoother" if omega == "chebyshev" else f"damped Jacobi omega {omega}"


def build_network(pm, torch, nel, levels, tol, precision, omega=0.6, steps=2, warm_net=None, coarse_space=True):
    from pymoto_amd import problems
    d = pm.VoxelDomain(*nel)
    bc, f = problems.cantilever3d(d)
    fd = torch.as_tensor(f, device="cuda")
    sx = pm.Signal("x", torch.full((d.nel,), 0.5, dtype=torch.float64, device="cuda"))
    if levels <= 0:
        levels = pm.solvers.auto_levels(d)
    mg = pm.solvers.build_hierarchy(d, levels, omega=omega, smooth_steps=steps, precision=precision)
    solver = pm.solvers.CG(preconditioner=mg, tol=tol)
    with pm.Network() as fn:
        sxf = pm.FilterConv(d, radius=1.5)(sx)
        ss = pm.MathExpression(f"{1e-9} + {1.0 - 1e-9}*inp0^3")(sxf)
        sK = pm.AssembleStiffness(d, bc=bc)(ss)
        inner = solver
        if warm_net is not None:
            from pymoto_amd.learned import WarmStartCG
            inner = WarmStartCG(solver, warm_net, d, sxf, coarse_space=coarse_space)
        lin = pm.LinSolve(hermitian=True, solver=inner)
        su = lin(sK, fd)
        sc = pm.EinSum("i,i->")(su, fd)
    return dict(domain=d, sx=sx, sc=sc, su=su, sK=sK, sxf=sxf, fn=fn, solver=solver, inner=inner, mg=mg,
                levels=levels, f=fd, lin=lin)


def train_warm_start(pm, torch, args, precision, omega, nsamples, epochs):
    """Train the learned warm start on solver-generated pairs: seeded designs from the benchmark's
    distribution (filtered uniform densities, seeds disjoint from the timed ones), solved with the
    given variant; returns (network, seconds spent on data and training, loss history)."""
    from pymoto_amd import learned
    t0 = time.perf_counter()
    net = build_network(pm, torch, tuple(args.nel), args.levels, args.tol, precision, omega=omega)
    nel = tuple(args.nel)
    samples = []
    for k in range(nsamples):
        x = torch.as_tensor(np.random.default_rng(5000 + k).uniform(0.0, 1.0, nel[0] * nel[1] * nel[2]),
                            device="cuda")
        design_iteration(net, x)
        samples.append((learned.make_input(net["domain"], net["sxf"].state, net["f"]),
                        learned.normalized_target(net["su"].state)))
    model = learned.WarmStartNet(nel, seed=0)
    hist = learned.train(model, samples, epochs=epochs, lr=2e-3)
    torch.cuda.synchronize()
    return model, time.perf_counter() - t0, hist


def learned_workload(pm, torch, nel=(128, 64, 64), pairs=2000, epochs=2, precision="fp32", omega=0.5, tol=1e-8,
                     n_eval=8, omega_data="chebyshev", maxit_data=3000):
    """Learned warm start / coarse correction at 128x64x64: ``pairs`` (input, displacement) pairs from
    seeded SIMP/OC trajectories solved by this package's MG-PCG (volume fractions 0.3/0.4/0.5, 60 design
    iterations each; Chebyshev smoother, at most ``maxit_data`` iterations per solve: late OC designs
    need 100-160 iterations at this size and Jacobi omega 0.5 occasionally many more), the network
    trained on them on the GPU, then evaluated on the designs of a held-out
    trajectory (another seed): CG iterations and seconds per solve for the plain MG-PCG, the network's warm
    start, and warm start plus the learned coarse correction (flexible PCG), every solve to tol; the
    compliance of each against an FP64 MG-PCG solve to 1e-12."""
    import time as _t
    from pymoto_amd import learned, problems
    t0 = _t.perf_counter()
    samples, info = learned.trajectory_samples(nel, pairs, seeds=(11, 12, 13), precision=precision,
                                               omega=omega_data, tol=tol, maxit=maxit_data)
    torch.cuda.synchronize()
    t_data = _t.perf_counter() - t0
    net = learned.WarmStartNet(nel, seed=0)
    t0 = _t.perf_counter()
    hist = learned.train(net, samples, epochs=epochs, lr=1e-3)
    torch.cuda.synchronize()
    t_train = _t.perf_counter() - t0
    del samples
    torch.cuda.empty_cache()
    # held-out trajectory: designs of another seed
    d = pm.VoxelDomain(*nel)
    bc, f = problems.cantilever3d(d)
    fd = torch.as_tensor(f, device="cuda")
    designs = []
    rng = np.random.default_rng(97)
    x = torch.as_tensor(np.clip(0.4 + 0.1 * (rng.random(d.nel) - 0.5), 0.0, 1.0), device="cuda")
    sx = pm.Signal("x", x)
    levels = pm.solvers.auto_levels(d)
    with pm.Network() as fn:
        sxf = pm.FilterConv(d, radius=1.5)(sx)
        ss = pm.MathExpression(f"{1e-9} + {1.0 - 1e-9}*inp0^3")(sxf)
        sK = pm.AssembleStiffness(d, bc=bc)(ss)
        su = pm.LinSolve(hermitian=True, solver=pm.solvers.CG(preconditioner=pm.solvers.build_hierarchy(
            d, levels, omega=omega, smooth_steps=2, precision=precision), tol=tol))(sK, fd)
        sc = pm.EinSum("i,i->")(su, fd)
    oc = pm.OC(sx, sc, fn, move=0.2, xmin=0.0, xmax=1.0, verbosity=0, maxvol=0.4)
    for it in range(n_eval):
        fn.reset()
        fn.response()
        designs.append((sxf.state.clone(), ss.state.clone()))
        sc.sensitivity = 1.0
        fn.sensitivity()
        sx.state = oc.step(sx.state, sx.sensitivity)
    out = {"plain": [], "warm": [], "warm+coarse": []}
    times = {k: 0.0 for k in out}
    comp = {k: [] for k in out}
    comp_ref = []
    for xf, xs in designs:
        K = pm.AssembleStiffness(d, bc=bc)(xs)
        ref = pm.solvers.CG(preconditioner=pm.solvers.build_hierarchy(d, levels, omega=omega, smooth_steps=2),
                            tol=1e-12, maxit=5000)
        ref.update(K)
        comp_ref.append(float(torch.dot(ref.solve(fd), fd)))
        for mode in out:
            cg = pm.solvers.CG(preconditioner=pm.solvers.build_hierarchy(d, levels, omega=omega, smooth_steps=2,
                                                                         precision=precision), tol=tol)
            solver = cg if mode == "plain" else learned.WarmStartCG(cg, net, d, pm.Signal("xf", xf),
                                                                   coarse_space=(mode == "warm+coarse"))
            solver.update(K)
            torch.cuda.synchronize()
            t1 = _t.perf_counter()
            u = solver.solve(fd)
            torch.cuda.synchronize()
            times[mode] += _t.perf_counter() - t1
            out[mode].append(int(solver.iterations))
            comp[mode].append(float(torch.dot(u, fd)))
    res = {"workload": f"learned warm start / coarse correction, cantilever {nel[0]}x{nel[1]}x{nel[2]} hex8, "
                       f"{precision} V-cycle (omega {omega}) in FP64 CG, tol {tol:g}",
           "training_pairs": info["pairs"], "trajectories": info["trajectories"], "epochs": epochs,
           "data_smoother": smoother_name(omega_data), "data_solver_iterations": info["solver_iterations"],
           "data_s": t_data, "train_s": t_train, "loss_history": [float(h) for h in hist],
           "params": int(net.nparams()), "held_out_designs": len(designs)}
    for mode in out:
        res[mode] = {"cg_iterations": out[mode], "s_per_solve": times[mode] / len(designs),
                     "compliance_rel_err_vs_fp64": float(max(abs(a - b) / abs(b) for a, b in zip(comp[mode],
                                                                                               comp_ref)))}
    return res


def design_iteration(net, x):
    net["sx"].state = x
    net["fn"].reset()
    net["fn"].response()
    net["sc"].sensitivity = 1.0
    net["fn"].sensitivity()
    return net["sc"].state


ROTATE = 4  # buffer sets cycled by the kernel timings: 4 x (u, y, densities) of 12.8M dofs = 955 MB in FP64


def kernel_roofline(pm, torch, nel, reps=20):
    """Average duration of the fine-grid K.u kernel (FP64, and the FP32 product of the V-cycle) from HIP
    events on the stream it is launched on, and the HBM roofline of its compulsory bytes: u read and
    y written (3 dofs per node) plus one density per element (the Dirichlet mask byte per node not
    counted).  The element product runs in the reflection basis (~0.24 kflop per element), so the
    kernel's roof is HBM, not the FP64 FMA rate.  Consecutive launches cycle through ROTATE sets of
    (densities, u, y): one set (239 MB in FP64) fits the 256 MB last-level cache, so repeated launches
    on one set would read it from there instead of HBM, as a K.u inside the solve (whose iteration
    touches ~2.4 GB) never does."""
    import json
    from pymoto_amd import _lib, problems
    d = pm.VoxelDomain(*nel)
    bc, _ = problems.cantilever3d(d)
    x = torch.as_tensor(np.random.default_rng(3).uniform(1e-3, 1.0, d.nel), device="cuda")
    K = pm.AssembleStiffness(d, bc=bc)(x)
    g = K.grid
    lib = _lib.lib()
    stream = torch.cuda.current_stream()
    sets64, sets32, keep = [], [], []
    for r in range(ROTATE):
        xr = (K.xe * (1.0 - 0.01 * r)).contiguous()
        u = torch.randn(g.n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(u)
        x32, u32 = xr.float(), u.float()
        y32 = torch.empty_like(u32)
        keep += [xr, u, y, x32, u32, y32]
        sets64.append((*g.nel3, _lib.ptr(g.K0_dev), g.k0_pattern, _lib.ptr(xr), _lib.ptr(g.bcmask_dev), g.bcdiag,
                       _lib.ptr(u), _lib.ptr(y), _lib.stream()))
        sets32.append((*g.nel3, _lib.ptr(g.K0_dev), g.k0_pattern, _lib.ptr(x32), _lib.ptr(g.bcmask_dev), g.bcdiag,
                       _lib.ptr(u32), _lib.ptr(y32), _lib.stream()))

    def timed(fn, sets):
        for q in range(len(sets)):
            _lib.check(fn(*sets[q]))
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for q in range(reps):
            _lib.check(fn(*sets[q % len(sets)]))
        e.record(stream)
        torch.cuda.synchronize()
        return s.elapsed_time(e) / reps * 1e-3

    t64 = timed(lib.pmt_stiffness_apply, sets64)
    t32 = timed(lib.pmt_stiffness_apply_f32, sets32)
    from pymoto_amd import tune
    tuning = {k: int(v) for k, v in tune.get_tuning().items()}
    b64 = 2.0 * g.n * 8 + d.nel * 8
    b32 = 2.0 * g.n * 4 + d.nel * 4
    out = dict(kernel="fine_refl_kernel<16, apply, double> (K.u, FP64, reflection basis)", seconds=t64, bytes=b64,
               gbs=b64 / t64 / 1e9, fp32=dict(kernel="fine_refl_kernel<16, apply, float>", us=t32 * 1e6,
                                                 bytes=b32, gbs=b32 / t32 / 1e9, frac=b32 / t32 / 1e9 / PEAK_HBM_GBS),
               tuning=tuning, traffic=None, rotating_sets=ROTATE)
    pmc = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "fine_refl_fp64_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as fh:
            rec = json.load(fh)
        if tuple(rec.get("nel", ())) == tuple(nel):
            out["traffic"] = rec["hbm_bytes_per_launch"]
    return out


def cg_iteration_roofline(pm, torch, nel, precision, omega, levels, tol=1e-8):
    """Whole-CG-iteration figure for a solver variant: milliseconds per PCG iteration of one cold solve on
    a seeded filtered design (HIP events around the solve on its stream, divided by the iterations), and
    the modeled algorithmic HBM bytes of one iteration - every kernel reading its inputs and writing its
    outputs once: the fine FP64 K.u fused with p.q and p.r (p, r, q, densities, mask), the r update (q, r
    read, r written), the V-cycle's first sweep from zero, the rest of its fine level (3 Jacobi sweeps
    and the residual: u, b, y, densities, mask each; restriction; prolongation), 1/7 more for the coarse
    levels (1/8 of the dofs per level), q.z, and the direction update with the x update (z, p, x read,
    p and x written).  Returns a dict."""
    from pymoto_amd import problems
    d = pm.VoxelDomain(*nel)
    bc, f = problems.cantilever3d(d)
    x = torch.as_tensor(np.random.default_rng(11).uniform(0.0, 1.0, d.nel), device="cuda")
    xf = pm.FilterConv(d, radius=1.5)(pm.Signal("x", x)).state
    xs = pm.MathExpression(f"{1e-9} + {1.0 - 1e-9}*inp0^3")(pm.Signal("xf", xf)).state
    K = pm.AssembleStiffness(d, bc=bc)(xs)
    cg = pm.solvers.CG(preconditioner=pm.solvers.build_hierarchy(d, levels, omega=omega, smooth_steps=2,
                                                                 precision=precision), tol=tol)
    cg.update(K)
    fd = torch.as_tensor(f, device="cuda")
    cg.solve(fd)  # warm-up (first-call costs)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    cg.solve(fd)
    e1.record(stream)
    torch.cuda.synchronize()
    its = int(cg.iterations)
    ms = e0.elapsed_time(e1) / max(its, 1)
    n, e, nn = 3 * d.nnodes, d.nel, d.nnodes
    S = 8.0
    s = 8.0 if precision == "fp64" else 4.0
    vfine = 4 * (3 * n * s + e * s + nn) + (n * s + n * s / 8) + (2 * n * s + n * s / 8)
    # the V-cycle's first sweep from zero (D^-1 and r read, u written) is fused into the x/r update for
    # reduced-precision V-cycles and a kernel of its own in FP64: 3 n s either way
    model = (3 * n * S + e * S + nn) + 3 * n * S + 3 * n * s + (8.0 / 7.0) * vfine + (n * s + 4 * n * S) \
        + (n * S + n * s)
    return {"variant": f"{precision}@{omega:g}", "cg_iterations": its, "ms_per_iteration": ms,
            "modeled_bytes": model, "achieved_gbs": model / (ms * 1e-3) / 1e9,
            "frac": model / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "note": "cold solve from zero on a seeded filtered design; bytes modeled, not counted"}


def ku_driver(pm, torch, nel, reps=5, which="all"):
    """Kernel-only run for rocprofv3 (--pmc counter passes, --kernel-trace): five launches each of the fine
    K.u in FP64 (reflection basis), FP32, BF16 (matrix cores), V-cycles (FP64, FP32), the FilterConv
    forward and adjoint, the SIMP power law forward and adjoint, the sensitivity contraction, the learned network's forward pass (conv3d on the matrix cores), then five
    plain copies of one nodal vector (calibration of the byte counters).  ``which`` = "ku" runs the FP64
    K.u and the copies only.  Prints one line per group."""
    from pymoto_amd import problems
    d = pm.VoxelDomain(*nel)
    bc, _ = problems.cantilever3d(d)
    x = torch.as_tensor(np.random.default_rng(0).uniform(1e-3, 1.0, d.nel), device="cuda")
    K = pm.AssembleStiffness(d, bc=bc)(x)
    u = torch.randn(3 * d.nnodes, dtype=torch.float64, device="cuda")
    # the K.u launches and the copies cycle through ROTATE buffer sets (none stays in the last-level cache)
    Ks = [K] + [pm.AssembleStiffness(d, bc=bc)(x * (1.0 - 0.01 * r)) for r in range(1, ROTATE)]
    us = [u] + [torch.randn_like(u) for _ in range(1, ROTATE)]
    for q in range(reps):
        Ks[q % ROTATE] @ us[q % ROTATE]
    torch.cuda.synchronize()
    print(f"fine K.u FP64 x{reps} ({ROTATE} rotating buffer sets)", flush=True)
    if which == "all":
        # V-cycles of the headline hierarchy (2 damped-Jacobi sweeps per level, omega 0.5): the fine sweeps
        # (fine_refl_kernel in Jacobi mode) in FP64 and FP32
        for prec in ("fp64", "fp32"):
            mg = pm.solvers.build_hierarchy(d, pm.solvers.auto_levels(d), omega=0.5, smooth_steps=2,
                                            precision=prec)
            mg.update(K)
            for _ in range(reps):
                mg.solve(u)
            torch.cuda.synchronize()
            print(f"V-cycle {prec} x{reps}", flush=True)
            del mg
        u32s = [v.to(torch.float32) for v in us]
        for prec in ("fp32", "bf16"):
            for q in range(reps):
                Ks[q % ROTATE].apply_reduced(u32s[q % ROTATE], prec)
            torch.cuda.synchronize()
            print(f"fine K.u {prec} x{reps}", flush=True)
        filt = pm.FilterConv(d, radius=1.5)
        sx = pm.Signal("x", x)
        sy = filt(sx)
        sy.sensitivity = torch.randn_like(x)
        for _ in range(reps):
            filt.response()
            sx.reset()
            filt.sensitivity()
        torch.cuda.synchronize()
        print(f"FilterConv forward + adjoint x{reps}", flush=True)
        simp = pm.MathExpression(f"{1e-9} + {1.0 - 1e-9}*inp0^3")
        sxp = pm.Signal("xp", x)
        syp = simp(sxp)
        syp.sensitivity = torch.randn_like(x)
        for _ in range(reps):
            simp.response()
            sxp.reset()
            simp.sensitivity()
        torch.cuda.synchronize()
        print(f"SIMP forward + adjoint x{reps}", flush=True)
        mod = pm.AssembleStiffness(d, bc=bc)
        sxs = pm.Signal("xs", x)
        sK = mod(sxs)
        v = torch.randn_like(u)
        for _ in range(reps):
            sK.sensitivity = pm.DyadicMatrix(u, v, fac=-1.0)
            sxs.reset()
            mod.sensitivity()
        torch.cuda.synchronize()
        print(f"sensitivity contraction x{reps}", flush=True)
        if all(n % 8 == 0 for n in nel):
            from pymoto_amd import learned
            net = learned.WarmStartNet(nel, seed=0)
            inp = learned.make_input(d, x, torch.zeros(3 * d.nnodes, dtype=torch.float64, device="cuda"))
            for _ in range(reps):
                net.forward(inp)
            torch.cuda.synchronize()
            print(f"learned network forward x{reps}", flush=True)
    cs = [torch.empty_like(u) for _ in range(ROTATE)]
    for q in range(reps):
        cs[q % ROTATE].copy_(us[q % ROTATE])
    torch.cuda.synchronize()
    print(f"copy x{reps}: vector of {u.numel()} FP64 = {u.numel() * 8 / 1024:.1f} KiB ({ROTATE} rotating pairs)",
          flush=True)


def batched_workload(pm, torch, nel=(128, 64, 64), beta=8.0, etas=(0.3, 0.5, 0.7), nloads=4, nuneven=16, seed=7,
                     precision="fp32", omega="chebyshev", omega_uneven=0.5, tol=1e-8, xmin=1e-9):
    """Robust multi-load batch: 3 Heaviside designs (eta 0.3/0.5/0.7, beta 8, SIMP xmin 1e-9: dilated /
    nominal / eroded projections as in the reference's robust example,
    examples/topology_optimization/ex_compliance_robust.py:137-149) x 4 seeded point loads on
    the x = max face, each design one 4-column block PCG (captured-graph iterations), one after the other
    and the three at once on three streams, timed against the same 12 solves one right-hand side at a time
    and as 12 concurrent single-right-hand-side solves (one handle and stream each); the fine operator's
    launch shape is picked by the autotuner first.  The robust designs use the Chebyshev smoother: the
    eroded design's lambda_max(D^-1 K) is about 4.7, so damped Jacobi with omega 0.5 amplifies its top
    modes (omega * lambda_max > 2) and PCG needs 2-6x the iterations or stalls.  Plus 16 independent
    problems of seeded sizes from 64x32 Q4 to 96x48x48 hex8 (filtered uniform densities, Jacobi
    omega_uneven), solved concurrently on 8 streams and timed against one stream."""
    import time as _t
    from pymoto_amd import problems
    from pymoto_amd.batch import solve_independent
    rng = np.random.default_rng(seed)
    d = pm.VoxelDomain(*nel)
    bc, _ = problems.cantilever3d(d)
    F = np.zeros((3 * d.nnodes, nloads))
    for c in range(nloads):
        j, k = int(rng.integers(0, nel[1] + 1)), int(rng.integers(0, nel[2] + 1))
        F[3 * d.get_nodenumber(nel[0], j, k) + 1, c] = -1.0
    Fd = torch.as_tensor(F, device="cuda")
    # densities in [0.35, 1]: after the r = 1.5 filter all three projections keep a connected structure
    # (from uniform[0, 1] the eroded design is almost void)
    x = torch.as_tensor(rng.uniform(0.35, 1.0, d.nel), device="cuda")
    xf = pm.FilterConv(d, radius=1.5)(pm.Signal("x", x)).state
    hv = "(tanh({b} * {e}) + tanh({b} * (inp0 - {e}))) / (tanh({b} * {e}) + tanh({b} * (1 - {e})))"
    Ks = []
    for eta in etas:
        xp = pm.MathExpression(hv.format(b=beta, e=eta))(pm.Signal("xf", xf)).state
        xs = pm.MathExpression(f"{xmin} + {1.0 - xmin}*inp0^3")(pm.Signal("xp", xp)).state
        Ks.append(pm.AssembleStiffness(d, bc=bc)(xs))
    levels = pm.solvers.auto_levels(d)

    def make_solver(graph=True):
        return pm.solvers.CG(preconditioner=pm.solvers.build_hierarchy(d, levels, omega=omega, smooth_steps=2,
                                                                       precision=precision), tol=tol, graph=graph,
                             maxit=2000)
    # launch shapes of the fine operator picked by the autotuner on this grid (applied process-wide)
    from pymoto_amd import tune
    tuned, _ = tune.autotune_fine_operator(Ks[0], precision=precision if precision != "fp64" else "fp64")
    solvers = [make_solver() for _ in etas]
    for K, sv in zip(Ks, solvers):  # setup + graph capture, untimed
        sv.update(K)
        sv.solve(Fd)
    torch.cuda.synchronize()
    t0 = _t.perf_counter()
    its_block = []
    for K, sv in zip(Ks, solvers):
        sv.update(K)
        sv.solve(Fd)
        its_block.append(int(sv.iterations))
    torch.cuda.synchronize()
    t_block = _t.perf_counter() - t0
    # the three block solves at once, one stream (and handle) per design
    solve_independent([(K, Fd, sv) for K, sv in zip(Ks, solvers)], workers=3)
    t0 = _t.perf_counter()
    _, its_block_conc = solve_independent([(K, Fd, sv) for K, sv in zip(Ks, solvers)], workers=3)
    t_block_conc = _t.perf_counter() - t0
    t0 = _t.perf_counter()
    its_single = []
    for K, sv in zip(Ks, solvers):
        sv.update(K)
        its_single.append([])
        for c in range(nloads):
            sv.solve(Fd[:, c].contiguous())
            its_single[-1].append(int(sv.iterations))
    torch.cuda.synchronize()
    t_single = _t.perf_counter() - t0
    # the 12 right-hand sides as 12 independent solves at once (a handle each, 12 streams)
    jobs12 = [(K, Fd[:, c].contiguous(), make_solver(graph=False)) for K in Ks for c in range(nloads)]
    solve_independent(jobs12, workers=12)  # setup, untimed
    t0 = _t.perf_counter()
    _, its_conc12 = solve_independent(jobs12, workers=12)
    t_single_conc = _t.perf_counter() - t0
    del jobs12

    # uneven batch of independent problems
    jobs = []
    dofs = 0
    for _ in range(nuneven):
        if rng.random() < 0.5:
            nx = int(rng.choice([64, 96, 128, 160, 192, 256]))
            dd = pm.VoxelDomain(nx, nx // 2)
            b2, f2 = problems.mbb2d(dd)
        else:
            nx = int(rng.choice([32, 48, 64, 80, 96]))
            dd = pm.VoxelDomain(nx, nx // 2, nx // 2)
            b2, f2 = problems.cantilever3d(dd)
        xx = torch.as_tensor(rng.uniform(0.0, 1.0, dd.nel), device="cuda")
        xxf = pm.FilterConv(dd, radius=1.5)(pm.Signal("x", xx)).state
        xxs = pm.MathExpression(f"{1e-9} + {1.0 - 1e-9}*inp0^3")(pm.Signal("xf", xxf)).state
        K2 = pm.AssembleStiffness(dd, bc=b2)(xxs)
        prec = precision if (dd.dim == 3 or precision != "bf16") else "fp32"
        sv = pm.solvers.CG(preconditioner=pm.solvers.build_hierarchy(dd, max(2, pm.solvers.auto_levels(dd)),
                                                                     omega=omega_uneven, smooth_steps=2,
                                                                     precision=prec),
                           tol=tol, maxit=2000)
        jobs.append((K2, torch.as_tensor(f2, device="cuda"), sv))
        dofs += K2.shape[0]
    solve_independent(jobs, workers=8)  # setup, untimed
    t0 = _t.perf_counter()
    _, its_uneven = solve_independent(jobs, workers=8)
    t_conc = _t.perf_counter() - t0
    t0 = _t.perf_counter()
    solve_independent(jobs, workers=1)
    t_seq = _t.perf_counter() - t0
    ndof = 3 * d.nnodes
    return {
        "workload": f"robust {nel[0]}x{nel[1]}x{nel[2]} hex8, Heaviside beta {beta} eta {list(etas)}, SIMP xmin {xmin:g}, "
                    f"{nloads} point loads per design; {precision} V-cycle ({smoother_name(omega)}), FP64 CG, "
                    f"tol {tol:g}",
        "rhs_solves": len(etas) * nloads, "ndof": ndof,
        "block_graph_s": t_block, "block_iterations": its_block,
        "block_graph_3_streams_s": t_block_conc, "block_3_streams_iterations": its_block_conc,
        "single_rhs_s": t_single, "single_rhs_iterations": its_single,
        "single_rhs_12_streams_s": t_single_conc, "single_rhs_12_streams_iterations": its_conc12,
        "dof_rhs_per_s_best": ndof * len(etas) * nloads / min(t_block, t_block_conc, t_single, t_single_conc),
        "autotuned_fine_operator": tuned,
        "uneven": {"problems": nuneven, "smoother": smoother_name(omega_uneven), "total_dofs": dofs,
                   "concurrent_8_streams_s": t_conc,
                   "one_stream_s": t_seq, "iterations": its_uneven},
    }


def cpu_baseline(nel, levels, omega, steps=3):
    """The CPU oracle (numpy/scipy restatement of the reference) on a bounded 3D sample, with the multigrid
    level rule and Jacobi damping of the GPU variant it is compared with.  Returns the mean seconds per
    design iteration (first iteration excluded when there are several), the sample's dof count and the CG
    iteration counts.  One core: scipy's sparse products and SuperLU are single-threaded, and the BLAS
    thread pools are limited to one thread for the measurement."""
    from threadpoolctl import threadpool_limits
    from oracle import problems
    from oracle.pipeline import design_iteration as cpu_iter
    bc, f = problems.cantilever3d(*nel)
    nelt = nel[0] * nel[1] * nel[2]
    times, its = [], []
    u = None
    with threadpool_limits(limits=1):
        for k in range(steps):
            x = np.random.default_rng(100 + k).uniform(0.0, 1.0, nelt)
            t0 = time.perf_counter()
            out = cpu_iter(tuple(nel), x, bc, f, radius=1.5, solver="mgcg", mg_levels=levels, omega=omega,
                           smooth_steps=2, tol=1e-8, x0=u)
            times.append(time.perf_counter() - t0)
            its.append(out["iterations"])
            u = out["u"]
    ndof = 3 * (nel[0] + 1) * (nel[1] + 1) * (nel[2] + 1)
    return float(np.mean(times[1:] if len(times) > 1 else times)), ndof, its


def run_variant(pm, torch, args, variant, xs, warm_net=None):
    precision, omega = parse_variant(variant)
    net = build_network(pm, torch, tuple(args.nel), args.levels, args.tol, precision,
                        omega=omega if omega is not None else args.omega, warm_net=warm_net)
    for k in range(args.warmup):
        design_iteration(net, xs[k])
    torch.cuda.synchronize()
    comp, its = [], []
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        comp.append(float(design_iteration(net, xs[k])))
        its.append(int(net["inner"].iterations))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    dcdx = net["sx"].sensitivity.clone()
    return net, dict(seconds=dt, cg_iterations=its, compliance=comp[-1], dcdx=dcdx,
                     mg_device_bytes=int(net["mg"].device_bytes()))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (the multi-GPU path is a design "
              f"note only)", file=sys.stderr)
        sys.exit(2)
    import torch
    import pymoto_amd as pm

    nel = tuple(args.nel)
    nelt = nel[0] * nel[1] * nel[2]
    if args.ku_only or args.kernels_only:
        ku_driver(pm, torch, nel, which="ku" if args.ku_only else "all")
        return
    nsteps = args.warmup + args.steps
    xs = [torch.as_tensor(np.random.default_rng(1000 + k).uniform(0.0, 1.0, nelt), device="cuda")
          for k in range(nsteps)]
    torch.cuda.synchronize()
    variants = [v.strip() for v in args.variants.split(",") if v.strip()]
    results, roof, levels = {}, None, None
    learned_info = None
    for v in variants:
        warm = None
        if v.endswith("+learned"):
            base = v[: -len("+learned")]
            prec, om = parse_variant(base)
            warm, t_train, hist = train_warm_start(pm, torch, args, prec, om if om is not None else args.omega,
                                                   args.learned_samples, args.learned_epochs)
            learned_info = dict(samples=args.learned_samples, epochs=args.learned_epochs, train_s=t_train,
                                loss_history=[float(h) for h in hist], params=int(warm.nparams()))
            v_run = base
        else:
            v_run = v
        net, res = run_variant(pm, torch, args, v_run, xs, warm_net=warm)
        results[v] = res
        levels = net["levels"]
        del net
        torch.cuda.empty_cache()
    roof = kernel_roofline(pm, torch, nel)
    ref = results.get("fp64")
    summary = {}
    for v, r in results.items():
        item = dict(s_per_iter=r["seconds"], cg_iterations=r["cg_iterations"], compliance=r["compliance"],
                    mg_device_bytes=r["mg_device_bytes"])
        if ref is not None and v != "fp64":
            item["compliance_rel_err_vs_fp64"] = abs(r["compliance"] - ref["compliance"]) / abs(ref["compliance"])
            item["dcdx_rel_err_vs_fp64"] = float((r["dcdx"] - ref["dcdx"]).abs().max() / ref["dcdx"].abs().max())
        summary[v] = item
    ok = [v for v in summary if v == "fp64" or ref is None or summary[v].get("compliance_rel_err_vs_fp64", 1.0) <= 1e-7]
    best = min(ok, key=lambda v: summary[v]["s_per_iter"])
    best_base = best.split("+")[0]
    same = "fp64" + (best_base[best_base.index("@"):] if "@" in best_base else "")
    best_omega = parse_variant(best)[1]
    best_omega = best_omega if isinstance(best_omega, float) else args.omega
    sec = summary[best]["s_per_iter"]
    d = pm.VoxelDomain(*nel)
    ndof = 3 * d.nnodes
    line = {
        "metric": "s/design-iter (solve+sens) @12.6M DOF: FP64 MG-PCG vs mixed+learned, res 1e-8",
        "value": sec,
        "unit": "s/design-iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f64 CG / f32 V-cycle",
                  "bf16": "f64 CG / f32 V-cycle, bf16 MFMA fine smoother"}[parse_variant(best)[0]],
        "data": "synthetic: seeded uniform[0,1] densities per step, filtered r=1.5",
        "config": {"workload": f"3D cantilever {nel[0]}x{nel[1]}x{nel[2]} hex8, MG-PCG ({levels} grids, V(2,2) "
                               f"damped Jacobi w={best_omega:g}, Galerkin coarse, dense direct coarsest), "
                               f"tol {args.tol:g}, warm start",
                   "ndof": ndof, "nel": d.nel, "variant": best, "dof_per_s": ndof / sec,
                   "value_is": "the fastest variant whose compliance agrees with the FP64 variant to 1e-7",
                   "variants": summary,
                   # headline speed-up: against FP64 with the same smoother (the precision gain alone)
                   "speedup_vs_fp64": (summary[same]["s_per_iter"] / sec) if same in summary else None,
                   "speedup_vs_fp64_reference_omega": (summary["fp64"]["s_per_iter"] / sec) if "fp64" in summary
                   else None},
    }
    learned_arms = [v for v in summary if v.endswith("+learned")]
    if learned_arms and best not in learned_arms:
        la = learned_arms[0]
        line["config"]["learned_arm_note"] = (
            f"the mixed+learned arm {la} takes {summary[la]['s_per_iter'] * 1e3:.1f} ms per design iteration with "
            f"CG iterations {summary[la]['cg_iterations']} against {sec * 1e3:.1f} ms and {summary[best]['cg_iterations']} "
            f"for {best}: the warm start and the deflation cost more than they save, so value is the mixed arm "
            f"without the network")
    if learned_info is not None:
        line["config"]["learned"] = learned_info
    if not args.no_batched:
        from pymoto_amd import tune
        before = tune.get_tuning()
        line["config"]["batched"] = batched_workload(pm, torch)
        tune.set_tuning(**before)  # the autotuner's shapes are for the batched grid, not for what follows
    if args.c5_pairs > 0:
        line["config"]["learned_128x64x64"] = learned_workload(pm, torch, pairs=args.c5_pairs, epochs=args.c5_epochs)
    bp, bo = parse_variant(best)
    line["config"]["cg_iteration_roofline"] = cg_iteration_roofline(
        pm, torch, nel, bp, bo if isinstance(bo, float) else args.omega, levels, args.tol)
    if roof is not None:
        line["roofline"] = {"bound": "hbm", "achieved": roof["gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": roof["gbs"] / PEAK_HBM_GBS, "traffic": roof["traffic"], "kernel": roof["kernel"],
                            "kernel_us": roof["seconds"] * 1e6, "algorithmic_bytes": roof["bytes"],
                            "fp32_product": roof["fp32"], "tuning": roof["tuning"],
                            "note": "compulsory bytes 2*ndof*8 + nel*8 per K.u; HIP events on the launch stream; "
                                    f"launches cycle through {ROTATE} buffer sets (none cache-resident)"}
    if not args.no_cpu:
        cpu_nel = tuple(args.cpu_nel)
        cpu_levels = pm.solvers.auto_levels(pm.VoxelDomain(*cpu_nel))
        best_prec, best_om = parse_variant(best)
        cpu_omega = best_om if isinstance(best_om, float) else args.omega
        t_cpu, ndof_cpu, its = cpu_baseline(cpu_nel, cpu_levels, cpu_omega)
        line["cpu_baseline"] = {
            "value": t_cpu * ndof / ndof_cpu, "unit": "s/design-iter", "cores": 1, "kind": "port",
            "scaled": True, "measured_value": t_cpu, "measured_ndof": ndof_cpu,
            "host_cpus": len(os.sched_getaffinity(0)),
            "thread_env": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS")},
            "sample": f"oracle (numpy/scipy restatement of pyMOTO, single-threaded scipy sparse + SuperLU) 3D "
                      f"cantilever {cpu_nel[0]}x{cpu_nel[1]}x{cpu_nel[2]} ({ndof_cpu} dofs), {cpu_levels}-grid MG-CG "
                      f"(the reference example's level rule) with omega {cpu_omega} as the GPU variant {best}, tol "
                      f"1e-8: {t_cpu:.3f} s/design-iter measured (CG its {its}); value = that time scaled linearly in "
                      f"dofs to {ndof} dofs (SCALED, not measured at that size)",
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
