"""Optimality-criteria design update on the device (``pymoto/common/optimizers.py:337-433``).

``OC.step`` follows the reference: positive sensitivities are clipped to zero, the move limits are
max(xmin, x - move) and min(xmax, x + move), and the Lagrange multiplier is bisected on
[l1init, l2init] until l2 - l1 <= l1l2tol, each trial design x*sqrt(-dg/lmid) clipped to the limits
and its volume compared with maxvol * n (maxvol defaults to the initial mean density).  Every trial
runs as one fused HIP kernel with a deterministic sum; the host reads 8 bytes per bisection step.
``OC.optimize`` drives a ``Network`` the way ``Optimizer.optimize`` does (response, sensitivity,
step, stopping on the relative design and objective change).
"""
import warnings

import numpy as np
import torch

from . import _lib


class OC:
    def __init__(self, variables, response, function, move=0.1, xmin=0.0, xmax=1.0, verbosity: int = 2,
                 l1init: float = 0.0, l2init: float = 100000.0, l1l2tol: float = 1e-4, maxvol: float = None):
        self.variable = variables[0] if isinstance(variables, (list, tuple)) else variables
        self.response = response
        self.function = function
        self.move, self.xmin, self.xmax = float(move), float(xmin), float(xmax)
        self.verbosity = verbosity
        self.l1init, self.l2init, self.l1l2tol = l1init, l2init, l1l2tol
        self.maxvol = maxvol
        self.iter = 0

    def step(self, x=None, dg=None):
        """New design from design x and objective sensitivity dg (device tensors)."""
        x = _lib.as_device(self.variable.state if x is None else x)
        dg = _lib.as_device(self.variable.sensitivity if dg is None else dg)
        n = x.numel()
        if self.maxvol is None:
            self.maxvol = float(x.sum().item()) / n
        if float(dg.max().item()) > 1e-15:
            warnings.warn("OC only works for negative sensitivities; clipping positive values.")
        lib = _lib.lib()
        xnew = torch.empty_like(x)
        tot = torch.empty(1, dtype=torch.float64, device=x.device)
        work = _lib.reduce_work()
        l1, l2 = self.l1init, self.l2init
        while l2 - l1 > self.l1l2tol:
            lmid = 0.5 * (l1 + l2)
            _lib.check(lib.pmt_oc_update(n, _lib.ptr(x), _lib.ptr(dg), self.xmin, self.xmax, self.move, lmid,
                                         _lib.ptr(xnew), _lib.ptr(tot), _lib.ptr(work), _lib.stream()), "OC update")
            if float(tot.item()) - self.maxvol * n > 0:
                l1 = lmid
            else:
                l2 = lmid
        return xnew

    def optimize(self, maxiter: int = 100, tolx: float = 1e-4, tolf: float = 1e-4, callback=None):
        """Iterate response -> sensitivity -> OC step; returns the objective history."""
        hist = []
        fprev = None
        for it in range(maxiter):
            self.function.reset()
            self.function.response()
            f = float(np.asarray(self.response.state))
            hist.append(f)
            self.response.sensitivity = 1.0
            self.function.sensitivity()
            x = _lib.as_device(self.variable.state)
            xnew = self.step(x, self.variable.sensitivity)
            if callback is not None:
                callback(it, x, xnew)
            dx = float((xnew - x).abs().max().item())
            self.variable.state = xnew
            self.iter = it + 1
            if self.verbosity >= 2:
                print(f"OC it {it}: f = {f:.6e}, max |dx| = {dx:.3e}")
            if dx < tolx or (fprev is not None and abs(f - fprev) <= tolf * abs(fprev)):
                break
            fprev = f
        return hist
