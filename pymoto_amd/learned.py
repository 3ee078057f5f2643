"""Learned warm start for the MG-PCG solve: a 3D convolutional encoder-decoder on the element grid.

The network maps the design (filtered density), the load (element-averaged nodal forces) and the
element coordinates to a displacement field; its prediction, scaled by the energy-optimal factor
alpha = (u.f) / (u.K u), is the initial guess of the CG solve.  It is trained on (design, load) ->
displacement pairs that the solver itself produces along optimisation trajectories.  Everything it
computes - convolutions (forward, backward to the input and to the weights), pooling, upsampling,
ReLU masks, the element-to-node map, the loss and the Adam update - runs in hand-written HIP
kernels of the library (``csrc/conv.hip``) with BF16 activations and weights and FP32 accumulation.

Architecture (levels halve the grid; widths 16/32/64/64; every conv 3x3x3, zero padding):

    in(8) -c16-c16-> s0 -pool-c32-c32-> s1 -pool-c64-c64-> s2 -pool-c64-c64- up+s2 -c64-c32- up+s1
          -c32-c16- up+s0 -c16-c16(linear)-> out (channels 0..2 averaged onto the nodes)

about 0.66 M parameters.  The input channels are: density, force x/y/z, coordinates x/y/z, 1.
``WarmStartCG`` wraps a ``CG`` solver: it starts from the better of the network's guess and the
caller's x0 (smaller residual), and falls back to the plain solve from the caller's x0 when the
guarded first iterations do not reduce the residual at least as fast as the plain start would
(either way the solve ends at the same FP64 residual tolerance).
"""
import ctypes

import numpy as np
import torch

from . import _lib

BF16 = torch.bfloat16

# (name, level, cin, cout, relu)
SPEC = [
    ("e0a", 0, 8, 16, True), ("e0b", 0, 16, 16, True),
    ("e1a", 1, 16, 32, True), ("e1b", 1, 32, 32, True),
    ("e2a", 2, 32, 64, True), ("e2b", 2, 64, 64, True),
    ("b0", 3, 64, 64, True), ("b1", 3, 64, 64, True),
    ("d2a", 2, 64, 64, True), ("d2b", 2, 64, 32, True),
    ("d1a", 1, 32, 32, True), ("d1b", 1, 32, 16, True),
    ("d0a", 0, 16, 16, True), ("d0b", 0, 16, 16, False),
]
NIN = 8


def _p(t):
    return _lib.ptr(t)


class WarmStartNet:
    """The encoder-decoder with FP32 master weights, BF16 copies and Adam state on the device."""

    def __init__(self, nel, seed: int = 0):
        nel = tuple(int(v) for v in nel)
        if len(nel) != 3 or any(n % 8 for n in nel):
            raise ValueError("the warm-start network needs a hex8 grid with element counts divisible by 8")
        self.nel = nel
        self.levels = [tuple(n >> l for n in nel) for l in range(4)]
        if any((a * b * c) % 16 for a, b, c in self.levels):
            raise ValueError("every level of the network needs a voxel count divisible by 16")
        dev = _lib.device()
        g = torch.Generator().manual_seed(seed)
        self.params = {}
        for name, _, cin, cout, relu in SPEC:
            std = np.sqrt((2.0 if relu else 1.0) / (27 * cin))
            if name == "d0b":
                std *= 0.1
            w = (torch.randn(27, cout, cin, generator=g, dtype=torch.float32) * std).to(dev)
            b = torch.zeros(cout, dtype=torch.float32, device=dev)
            self.params[name] = dict(w=w, b=b, wb=w.to(BF16), mw=torch.zeros_like(w), vw=torch.zeros_like(w),
                                     mb=torch.zeros_like(b), vb=torch.zeros_like(b), gw=torch.zeros_like(w),
                                     gb=torch.zeros_like(b))
        self.t = 0
        self.lib = _lib.lib()
        maxw = max(27 * cin * cout for _, _, cin, cout, _ in SPEC)
        self._wt = torch.empty(maxw, dtype=BF16, device=dev)
        self._work = torch.empty(max(self.lib.pmt_conv3d_wgrad_work_size(cin, cout) for _, _, cin, cout, _ in SPEC),
                                 dtype=torch.float32, device=dev)
        self.acts = {}

    def nparams(self):
        return sum(p["w"].numel() + p["b"].numel() for p in self.params.values())

    # ------------------------------------------------------------------ building blocks
    def _conv(self, name, x):
        _, lvl, cin, cout, relu = next(s for s in SPEC if s[0] == name)
        nx, ny, nz = self.levels[lvl]
        y = torch.empty(nx * ny * nz, cout, dtype=BF16, device=x.device)
        p = self.params[name]
        _lib.check(self.lib.pmt_conv3d_fwd(nx, ny, nz, cin, cout, _p(x), _p(p["wb"]), _p(p["b"]), int(relu), _p(y),
                                           _lib.stream()), f"conv {name}")
        return y

    def _pool(self, lvl, x, c):
        nx, ny, nz = self.levels[lvl]
        y = torch.empty((nx // 2) * (ny // 2) * (nz // 2), c, dtype=BF16, device=x.device)
        _lib.check(self.lib.pmt_pool2_bf16(nx, ny, nz, c, _p(x), _p(y), _lib.stream()), "pool")
        return y

    def _up_add(self, lvl, x, skip, c):
        nx, ny, nz = self.levels[lvl]
        y = torch.empty_like(skip)
        _lib.check(self.lib.pmt_up2_add_bf16(nx, ny, nz, c, _p(x), _p(skip), _p(y), _lib.stream()), "upsample")
        return y

    # ------------------------------------------------------------------ forward / backward
    def forward(self, inp):
        """inp: (nel, 8) BF16 channels-last.  Returns the nodal prediction (FP32, 3*nnodes)."""
        a = self.acts
        a["in"] = inp
        a["e0a"] = self._conv("e0a", inp)
        a["e0b"] = self._conv("e0b", a["e0a"])
        a["p0"] = self._pool(0, a["e0b"], 16)
        a["e1a"] = self._conv("e1a", a["p0"])
        a["e1b"] = self._conv("e1b", a["e1a"])
        a["p1"] = self._pool(1, a["e1b"], 32)
        a["e2a"] = self._conv("e2a", a["p1"])
        a["e2b"] = self._conv("e2b", a["e2a"])
        a["p2"] = self._pool(2, a["e2b"], 64)
        a["b0"] = self._conv("b0", a["p2"])
        a["b1"] = self._conv("b1", a["b0"])
        a["u2"] = self._up_add(2, a["b1"], a["e2b"], 64)
        a["d2a"] = self._conv("d2a", a["u2"])
        a["d2b"] = self._conv("d2b", a["d2a"])
        a["u1"] = self._up_add(1, a["d2b"], a["e1b"], 32)
        a["d1a"] = self._conv("d1a", a["u1"])
        a["d1b"] = self._conv("d1b", a["d1a"])
        a["u0"] = self._up_add(0, a["d1b"], a["e0b"], 16)
        a["d0a"] = self._conv("d0a", a["u0"])
        a["out"] = self._conv("d0b", a["d0a"])
        ex, ey, ez = self.nel
        u = torch.empty(3 * (ex + 1) * (ey + 1) * (ez + 1), dtype=torch.float32, device=inp.device)
        _lib.check(self.lib.pmt_elem_to_node(ex, ey, ez, 16, _p(a["out"]), _p(u), _lib.stream()), "elem to node")
        return u

    def _conv_bwd(self, name, x_in, g_out, need_dx=True):
        """Weight/bias gradients of layer `name` (g_out already masked by its ReLU) and dX."""
        _, lvl, cin, cout, _ = next(s for s in SPEC if s[0] == name)
        nx, ny, nz = self.levels[lvl]
        p = self.params[name]
        _lib.check(self.lib.pmt_conv3d_bwd_weight(nx, ny, nz, cin, cout, _p(x_in), _p(g_out), _p(p["gw"]),
                                                  _p(p["gb"]), _p(self._work), _lib.stream()), f"wgrad {name}")
        if not need_dx:
            return None
        dx = torch.empty(nx * ny * nz, cin, dtype=BF16, device=g_out.device)
        _lib.check(self.lib.pmt_conv3d_bwd_data(nx, ny, nz, cin, cout, _p(g_out), _p(p["wb"]), _p(self._wt), _p(dx),
                                                _lib.stream()), f"dgrad {name}")
        return dx

    def _mask(self, y, g):
        _lib.check(self.lib.pmt_relu_mask_bf16(g.numel(), _p(y), _p(g), _lib.stream()), "relu mask")
        return g

    def _add(self, a, b):
        y = torch.empty_like(a)
        _lib.check(self.lib.pmt_add_bf16(a.numel(), _p(a), _p(b), _p(y), _lib.stream()), "add")
        return y

    def _up_bwd(self, lvl, g, c):
        nx, ny, nz = self.levels[lvl]
        d = torch.empty((nx // 2) * (ny // 2) * (nz // 2), c, dtype=BF16, device=g.device)
        _lib.check(self.lib.pmt_up2_bwd_bf16(nx, ny, nz, c, _p(g), _p(d), _lib.stream()), "upsample adjoint")
        return d

    def _pool_bwd(self, lvl, g, c):
        nx, ny, nz = self.levels[lvl]
        d = torch.empty(nx * ny * nz, c, dtype=BF16, device=g.device)
        _lib.check(self.lib.pmt_pool2_bwd_bf16(nx, ny, nz, c, _p(g), _p(d), _lib.stream()), "pool adjoint")
        return d

    def backward(self, du):
        """Gradients of all weights for the nodal output gradient du (FP32)."""
        a = self.acts
        ex, ey, ez = self.nel
        g = torch.empty(ex * ey * ez, 16, dtype=BF16, device=du.device)
        _lib.check(self.lib.pmt_elem_to_node_bwd(ex, ey, ez, 16, _p(du), _p(g), _lib.stream()), "elem to node adjoint")
        g = self._conv_bwd("d0b", a["d0a"], g)
        g = self._conv_bwd("d0a", a["u0"], self._mask(a["d0a"], g))
        g_s0 = g  # up+skip: the skip receives g unchanged
        g = self._up_bwd(0, g, 16)
        g = self._conv_bwd("d1b", a["d1a"], self._mask(a["d1b"], g))
        g = self._conv_bwd("d1a", a["u1"], self._mask(a["d1a"], g))
        g_s1 = g
        g = self._up_bwd(1, g, 32)
        g = self._conv_bwd("d2b", a["d2a"], self._mask(a["d2b"], g))
        g = self._conv_bwd("d2a", a["u2"], self._mask(a["d2a"], g))
        g_s2 = g
        g = self._up_bwd(2, g, 64)
        g = self._conv_bwd("b1", a["b0"], self._mask(a["b1"], g))
        g = self._conv_bwd("b0", a["p2"], self._mask(a["b0"], g))
        g = self._add(self._pool_bwd(2, g, 64), g_s2)
        g = self._conv_bwd("e2b", a["e2a"], self._mask(a["e2b"], g))
        g = self._conv_bwd("e2a", a["p1"], self._mask(a["e2a"], g))
        g = self._add(self._pool_bwd(1, g, 32), g_s1)
        g = self._conv_bwd("e1b", a["e1a"], self._mask(a["e1b"], g))
        g = self._conv_bwd("e1a", a["p0"], self._mask(a["e1a"], g))
        g = self._add(self._pool_bwd(0, g, 16), g_s0)
        g = self._conv_bwd("e0b", a["e0a"], self._mask(a["e0b"], g))
        self._conv_bwd("e0a", a["in"], self._mask(a["e0a"], g), need_dx=False)

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        for p in self.params.values():
            _lib.check(self.lib.pmt_adam_step(p["w"].numel(), lr, b1, b2, eps, self.t, _p(p["gw"]), _p(p["w"]),
                                              _p(p["mw"]), _p(p["vw"]), _p(p["wb"]), _lib.stream()), "adam")
            _lib.check(self.lib.pmt_adam_step(p["b"].numel(), lr, b1, b2, eps, self.t, _p(p["gb"]), _p(p["b"]),
                                              _p(p["mb"]), _p(p["vb"]), None, _lib.stream()), "adam")


# ---------------------------------------------------------------------- data, loss, training
def make_input(domain, xfilt, f):
    """(nel, 8) BF16 input: density, element-averaged force (x, y, z, scaled to max 1), coordinates, 1."""
    ex, ey, ez = domain.nelx, domain.nely, domain.nelz
    dev = _lib.device()
    x = _lib.as_device(xfilt).to(torch.float32)
    F = _lib.as_device(f).to(torch.float32).reshape(ez + 1, ey + 1, ex + 1, 3)
    Fe = (F[:-1, :-1, :-1] + F[:-1, :-1, 1:] + F[:-1, 1:, :-1] + F[:-1, 1:, 1:] + F[1:, :-1, :-1] + F[1:, :-1, 1:]
          + F[1:, 1:, :-1] + F[1:, 1:, 1:]) / 8.0
    Fe = Fe.reshape(-1, 3)
    Fe = Fe / Fe.abs().max().clamp_min(1e-30)
    k, j, i = torch.meshgrid(torch.arange(ez, device=dev), torch.arange(ey, device=dev), torch.arange(ex, device=dev),
                             indexing="ij")
    coords = torch.stack([(i.reshape(-1) + 0.5) / ex, (j.reshape(-1) + 0.5) / ey, (k.reshape(-1) + 0.5) / ez], 1)
    inp = torch.cat([x[:, None], Fe, coords.float(), torch.ones_like(x)[:, None]], 1)
    return inp.to(BF16).contiguous()


def _lib_dot(a, b):
    a64 = a.to(torch.float64)
    b64 = b.to(torch.float64)
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    _lib.check(_lib.lib().pmt_dot(a64.numel(), _p(a64), _p(b64), _p(out), _p(_lib.reduce_work()), _lib.stream()),
               "dot")
    return float(out.item())


def normalized_target(u):
    """u / rms(u) in FP32 (the network learns the shape; the solve picks the scale)."""
    u = _lib.as_device(u)
    rms = np.sqrt(_lib_dot(u, u) / u.numel())
    return (u / (rms if rms > 0 else 1.0)).to(torch.float32).contiguous()


def loss_and_grad(pred, target):
    """Mean squared error over the nodal dofs and its gradient (FP32)."""
    diff = (pred - target).contiguous()
    n = diff.numel()
    return _lib_dot(diff, diff) / n, (diff * (2.0 / n)).contiguous()


def train(net, samples, epochs=3, lr=1e-3, seed=0, verbose=False):
    """samples: list of (input BF16 (nel,8), normalised target FP32).  Returns the loss history."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for idx in order:
            inp, tgt = samples[idx]
            pred = net.forward(inp)
            loss, g = loss_and_grad(pred, tgt)
            net.backward(g)
            net.adam_step(lr=lr)
            tot += loss
        hist.append(tot / len(samples))
        if verbose:
            print(f"epoch {ep}: mean loss {hist[-1]:.4e}", flush=True)
    return hist


def predict_displacement(net, domain, K, xfilt, f):
    """Warm start alpha * u_net with alpha = (u.f)/(u.K u) (energy-optimal scale), Dirichlet dofs zero."""
    u = net.forward(make_input(domain, xfilt, f)).to(torch.float64)
    g = K.grid
    mask = torch.zeros_like(u)
    _lib.check(_lib.lib().pmt_bc_split(*g.nel3, _p(g.bcmask_dev), 1.0, _p(u), _p(mask), _p(torch.empty_like(u)),
                                       _lib.stream()), "bc split")
    u = mask
    fd = _lib.as_device(f)
    uku = _lib_dot(u, K @ u)
    alpha = _lib_dot(u, fd) / uku if uku > 0 else 0.0
    return alpha * u


def coarse_space(A, u):
    """The learned coarse space of one solve: W = the prediction u split into its three displacement
    components (three FP64 columns, Dirichlet dofs zero), AW = A W and Einv = (W^T A W)^-1 (host).
    Columns whose energy vanishes are dropped.  Returns (W (m, n), AW (m, n), Einv (m, m) numpy)."""
    u = _lib.as_device(u)
    g = A.grid
    cols = []
    for d in range(3):
        w = torch.zeros_like(u)
        w[d::3] = u[d::3]  # one displacement component (a strided copy)
        wm = torch.empty_like(w)
        _lib.check(_lib.lib().pmt_bc_split(*g.nel3, _p(g.bcmask_dev), 1.0, _p(w), _p(wm), _p(torch.empty_like(w)),
                                           _lib.stream()), "bc split")
        cols.append(wm)
    AWl = [A @ w for w in cols]
    E = np.array([[_lib_dot(cols[i], AWl[j]) for j in range(3)] for i in range(3)])
    keep = [i for i in range(3) if E[i, i] > 1e-12 * max(np.max(np.diag(E)), 1e-300)]
    W = torch.stack([cols[i] for i in keep]).contiguous()
    AW = torch.stack([AWl[i] for i in keep]).contiguous()
    Einv = np.linalg.inv(0.5 * (E[np.ix_(keep, keep)] + E[np.ix_(keep, keep)].T))
    return W, AW, Einv


def deflated_start(A, W, Einv, b, x0=None):
    """x0 + W Einv W^T (b - A x0): the start whose residual is orthogonal to span W (the Galerkin projection
    of the solution onto W when x0 = 0), which deflated CG needs."""
    b = _lib.as_device(b)
    x = torch.zeros_like(b) if x0 is None else _lib.as_device(x0).clone()
    r = A.residual(b, x)
    coef = Einv @ np.array([_lib_dot(W[i], r) for i in range(W.shape[0])])
    for i in range(W.shape[0]):
        _lib.check(_lib.lib().pmt_axpby(x.numel(), float(coef[i]), _p(W[i]), 1.0, _p(x), _lib.stream()), "axpby")
    return x


class WarmStartCG:
    """CG whose initial guess comes from the network when that guess has the smaller residual;
    a guarded start falls back to the plain solve (from the caller's x0) if the network start stalls.

    ``coarse_space=True`` (PCG with a learned coarse correction): the prediction's three displacement
    components span a coarse space W; the solve starts from the caller's start corrected by
    W (W^T A W)^-1 W^T r0 (its residual orthogonal to W) and every preconditioned residual is projected
    A-orthogonally to W (``pmt_mg_set_coarse_space``: deflated PCG, whose directions never revisit the
    error components that W already resolves)."""

    def __init__(self, cg, net, domain, xfilt_signal, guard_its: int = 5, coarse_space: bool = False):
        self.cg, self.net, self.domain, self.xsig = cg, net, domain, xfilt_signal
        self.guard_its = guard_its
        self.coarse = coarse_space
        self.used_network = False
        self.iterations = 0
        self.A = None

    def update(self, A):
        self.A = A
        self.cg.update(A)

    def _solve_coarse(self, rhs, x0):
        u = self.net.forward(make_input(self.domain, self.xsig.state, rhs)).to(torch.float64)
        W, AW, Einv = coarse_space(self.A, u)
        start = deflated_start(self.A, W, Einv, rhs, x0)
        self.used_network = True
        lib = _lib.lib()
        h = self.cg.preconditioner.handle
        Eh = np.ascontiguousarray(Einv, dtype=np.float64)
        _lib.check(lib.pmt_mg_set_coarse_space(h, W.shape[0], _p(W), _p(AW), Eh.ctypes.data_as(ctypes.c_void_p)),
                   "coarse space")
        try:
            out = self.cg.solve(rhs, x0=start)
        finally:
            _lib.check(lib.pmt_mg_set_coarse_space(h, 0, None, None, None), "coarse space")
        self.iterations = self.cg.iterations
        return out

    def solve(self, rhs, x0=None, trans="N"):
        rhs = _lib.as_device(rhs)
        if rhs.ndim != 1:
            return self.cg.solve(rhs, x0=x0, trans=trans)
        if self.coarse:
            return self._solve_coarse(rhs, x0)
        guess = predict_displacement(self.net, self.domain, self.A, self.xsig.state, rhs)
        rg = self.A.residual(rhs, guess)
        r_net = np.sqrt(_lib_dot(rg, rg))
        if x0 is not None:
            r0 = self.A.residual(rhs, _lib.as_device(x0))
            r_plain = np.sqrt(_lib_dot(r0, r0))
        else:
            r_plain = np.sqrt(_lib_dot(rhs, rhs))
        self.used_network = r_net < r_plain
        if not self.used_network:
            out = self.cg.solve(rhs, x0=x0)
            self.iterations = self.cg.iterations
            return out
        # guarded start: a few iterations from the network guess must beat the plain start's residual
        maxit = self.cg.maxit
        self.cg.maxit = self.guard_its
        try:
            import warnings
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                xg = self.cg.solve(rhs, x0=guess)
            its = self.cg.iterations
        finally:
            self.cg.maxit = maxit
        if self.cg.relres <= self.cg.tol:
            self.iterations = its
            return xg
        stall = self.cg.relres * np.sqrt(_lib_dot(rhs, rhs)) > r_plain
        if stall:
            self.used_network = False
            out = self.cg.solve(rhs, x0=x0)
            self.iterations = its + self.cg.iterations
            return out
        out = self.cg.solve(rhs, x0=xg)
        self.iterations = its + self.cg.iterations
        return out


# ---------------------------------------------------------------------- solver-generated training data
def trajectory_samples(nel, n_pairs, seeds=(0,), volfracs=(0.3, 0.4, 0.5), iters=60, precision="fp32", omega=0.5,
                       tol=1e-8, radius=1.5, move=0.2, keep_every=1, maxit=10000):
    """(input, normalised target) pairs from seeded SIMP/OC trajectories of the cantilever (filter r=1.5,
    SIMP p=3, the MG-PCG solve of this package, OC with the reference's bisection,
    ``pymoto/common/optimizers.py:337-437``): trajectory t starts from the uniform design at volume
    fraction volfracs[t % len] plus seeded noise, and every ``keep_every``-th design iteration gives one
    pair, until ``n_pairs`` pairs exist; ``omega`` is the Jacobi factor or "chebyshev", ``maxit`` bounds each
    solve.  Returns (samples, info)."""
    import pymoto_amd as pm
    from . import problems
    d = pm.VoxelDomain(*nel)
    bc, f = problems.cantilever3d(d)
    fd = _lib.as_device(f)
    levels = pm.solvers.auto_levels(d)
    samples, traj, total_its = [], 0, 0
    while len(samples) < n_pairs:
        seed = seeds[traj % len(seeds)] + 1000 * (traj // len(seeds))
        vf = volfracs[traj % len(volfracs)]
        rng = np.random.default_rng(seed)
        x0 = np.clip(vf + 0.1 * (rng.random(d.nel) - 0.5), 0.0, 1.0)
        sx = pm.Signal("x", _lib.as_device(x0))
        solver = pm.solvers.CG(preconditioner=pm.solvers.build_hierarchy(d, levels, omega=omega, smooth_steps=2,
                                                                         precision=precision), tol=tol,
                                   maxit=maxit)
        with pm.Network() as fn:
            sxf = pm.FilterConv(d, radius=radius)(sx)
            ss = pm.MathExpression(f"{1e-9} + {1.0 - 1e-9}*inp0^3")(sxf)
            sK = pm.AssembleStiffness(d, bc=bc)(ss)
            su = pm.LinSolve(hermitian=True, solver=solver)(sK, fd)
            sc = pm.EinSum("i,i->")(su, fd)
        oc = pm.OC(sx, sc, fn, move=move, xmin=0.0, xmax=1.0, verbosity=0, maxvol=vf)
        for it in range(iters):
            fn.reset()
            fn.response()
            total_its += solver.iterations
            if it % keep_every == 0:
                samples.append((make_input(d, sxf.state, fd), normalized_target(su.state)))
                if len(samples) >= n_pairs:
                    break
            sc.sensitivity = 1.0
            fn.sensitivity()
            sx.state = oc.step(sx.state, sx.sensitivity)
        traj += 1
    return samples, dict(trajectories=traj, pairs=len(samples), solver_iterations=total_its)
