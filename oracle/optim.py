"""Oracle of the OC design update (test infrastructure only).

Restates ``pymoto/common/optimizers.py:404-433`` (OC.step): clip positive sensitivities, move
limits max(xmin, x - move) / min(xmax, x + move), bisection of the multiplier on [l1init, l2init]
until l2 - l1 <= l1l2tol with the volume test sum(xnew) - maxvol * n > 0.

Pinning: tests/test_oracle.py checks that the result meets the volume constraint to the bisection
tolerance, respects the move limits and equals the closed-form update for a uniform sensitivity.
"""
import numpy as np


def oc_step(x, dg, maxvol=None, move=0.1, xmin=0.0, xmax=1.0, l1init=0.0, l2init=100000.0, l1l2tol=1e-4):
    if maxvol is None:
        maxvol = np.sum(x) / x.size
    dg = np.minimum(dg, 0)
    lb = np.maximum(xmin, x - move)
    ub = np.minimum(xmax, x + move)
    xnew = x.copy()
    l1, l2 = l1init, l2init
    while l2 - l1 > l1l2tol:
        lmid = 0.5 * (l1 + l2)
        xnew[:] = np.clip(x * np.sqrt(-dg / lmid), lb, ub)
        l1, l2 = (lmid, l2) if np.sum(xnew) - maxvol * x.size > 0 else (l1, lmid)
    return xnew
