"""Autotuner for the launch shapes of the fine-grid operator.

The fine K u kernel dominates every multigrid-preconditioned solve; its best launch shape depends
on the grid extents (how the node columns tile the x-y plane, how many z planes a block streams)
and on the device.  ``autotune_fine_operator`` times the candidate shapes on the actual operator
with HIP events on the launch stream, keeps the fastest, and applies it process-wide through
``pmt_set_tuning`` (read by every later launch, including captured graphs, which are re-captured).
Results are cached per (grid, device) for the life of the process.

Tuning parameters (``include/pymoto_amd.h``): ``FINE_KERNEL`` (0 = 3D tiles, 1 = z-streamed
columns, 2 = z-streamed columns multiplying in the reflection basis of K0), ``STREAM_TILE`` (0 = 32x8,
1 = 64x4, 2 = 16x16 nodes per plane of kernel 1), ``REFL_ROWS`` (block shape of kernel 2: 0 = chosen
per grid, 16 / 24 / 32 rows of 16 nodes, 12 rows of 32 nodes), ``STREAM_PLANES`` (z planes per block, 0 = derived from the CU count).
"""
import itertools

import torch

from . import _lib

FINE_KERNEL, STREAM_PLANES, STREAM_TILE, REFL_ROWS, COARSE1 = 1, 2, 3, 4, 5
DEFAULT = dict(fine_kernel=2, stream_tile=0, stream_planes=0, refl_rows=16)
_cache = {}


def get_tuning():
    lib = _lib.load()
    return {"fine_kernel": lib.pmt_get_tuning(FINE_KERNEL), "stream_tile": lib.pmt_get_tuning(STREAM_TILE),
            "stream_planes": lib.pmt_get_tuning(STREAM_PLANES), "refl_rows": lib.pmt_get_tuning(REFL_ROWS)}


def set_tuning(fine_kernel=None, stream_tile=None, stream_planes=None, refl_rows=None):
    lib = _lib.lib()
    for param, value in ((FINE_KERNEL, fine_kernel), (STREAM_TILE, stream_tile), (STREAM_PLANES, stream_planes),
                         (REFL_ROWS, refl_rows)):
        if value is not None:
            _lib.check(lib.pmt_set_tuning(param, int(value)), "set tuning")


def set_coarse_level1(stencil_free=True):
    """Operator of the first coarse level of hex8 hierarchies created from now on: element by element in
    the reflection basis of K0 (default) or the stored half stencil."""
    _lib.check(_lib.lib().pmt_set_tuning(COARSE1, 1 if stencil_free else 0), "set tuning")


def candidates(nz, reflection=True):
    """Launch shapes tried for a grid with nz node planes (kernel 2 only when K0 has the
    reflection form)."""
    out = [dict(fine_kernel=0, stream_tile=0, stream_planes=0, refl_rows=16)]
    planes = sorted({0, 4, 8, 16, max(2, nz // 4)})
    for tile, p in itertools.product((0, 1, 2), planes):
        out.append(dict(fine_kernel=1, stream_tile=tile, stream_planes=p, refl_rows=16))
    if reflection:
        for rows, p in itertools.product((0, 12, 16, 24, 32), planes):
            out.append(dict(fine_kernel=2, stream_tile=0, stream_planes=p, refl_rows=rows))
    return out


def _time(fn, reps):
    fn()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps * 1e-3


def autotune_fine_operator(K, reps: int = 10, precision: str = "fp64", verbose: bool = False):
    """Pick the fastest launch shape for ``K @ u`` (``precision`` 'fp64', or 'fp32'/'bf16' for the
    reduced-precision sweeps) and apply it.  Returns (best shape, {shape: seconds})."""
    g = K.grid
    dev = torch.cuda.get_device_name()
    key = (g.nel3, precision, dev)
    if key in _cache:
        best, table = _cache[key]
        set_tuning(**best)
        return best, table
    u = torch.randn(g.n, dtype=torch.float64, device=K.xe.device)
    if precision == "fp64":
        def fn():
            K @ u
    else:
        u32 = u.to(torch.float32)

        def fn():
            K.apply_reduced(u32, precision)
    before = get_tuning()
    table = {}
    try:
        for c in candidates(g.nel3[2] + 1 if g.nel3[2] else 1, reflection=bool(g.k0_pattern & 2)):
            if g.nel3[2] == 0 and c["fine_kernel"] != 0:
                continue  # 2D grids use the tiled kernel
            if precision == "bf16" and c["fine_kernel"] == 0:
                continue  # the matrix-core operator is always streamed
            set_tuning(**c)
            table[tuple(sorted(c.items()))] = _time(fn, reps)
    except Exception:
        set_tuning(**before)
        raise
    best = dict(min(table, key=table.get))
    set_tuning(**best)
    if verbose:
        for k, t in sorted(table.items(), key=lambda kv: kv[1]):
            print(f"  {dict(k)}: {t * 1e6:.1f} us")
    _cache[key] = (best, table)
    return best, table
