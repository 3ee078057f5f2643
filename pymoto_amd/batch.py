"""Batched solves: many right-hand sides of one operator, and many independent problems at once.

* Several load cases of one stiffness matrix go through ``CG.solve`` with an (ndof, k) right-hand
  side: one block PCG (``pymoto/solvers/iterative.py:340-403``) whose iterations replay captured
  HIP graphs (``pmt_block_pcg_solve``).
* Independent problems (different grids, designs or loads) are solved concurrently by
  ``solve_independent``: each runs on its own HIP stream, driven by its own host thread (the
  library releases the Python lock during every call and keeps its per-problem state in the
  multigrid handle), so the small problems' kernels overlap on the device instead of queueing
  behind one another.
"""
import threading
from concurrent.futures import ThreadPoolExecutor

import torch

from . import _lib


def solve_independent(jobs, workers: int = 8):
    """Solve independent systems concurrently.

    ``jobs``: list of (K, f, solver) with K a device ``StiffnessMatrix``, f a device right-hand
    side (ndof,) or (ndof, k) and solver a ``CG`` with a ``GeometricMultigrid`` preconditioner
    (not shared between jobs).  Returns the list of solutions and the list of iteration counts.
    """
    local = threading.local()
    dev = _lib.device()

    def run(job):
        K, f, solver = job
        if not hasattr(local, "stream"):
            local.stream = torch.cuda.Stream(device=dev)
        with torch.cuda.stream(local.stream):
            solver.update(K)
            u = solver.solve(f)
        local.stream.synchronize()
        return u, solver.iterations

    torch.cuda.synchronize()
    if workers <= 1 or len(jobs) <= 1:
        out = [run(j) for j in jobs]
    else:
        with ThreadPoolExecutor(max_workers=min(workers, len(jobs))) as ex:
            out = list(ex.map(run, jobs))
    torch.cuda.synchronize()
    return [o[0] for o in out], [o[1] for o in out]
