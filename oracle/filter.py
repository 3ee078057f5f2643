"""Oracle: FilterConv forward and adjoint (test infrastructure only).

Restates ``pymoto/modules/filter.py:8-220``: weights ``max(0, r - |d|)`` on a box of half-width
``floor((r - 1e-10 h)/h)`` clipped to the domain, normalised to sum 1 (``filter.py:189-205``);
padding of the element-index array per face with ``np.pad`` modes 'symmetric'/'edge'/'wrap' or a
constant via overrides (``filter.py:98-187``); forward ``scipy.signal.convolve(..., 'valid')``
(``filter.py:207-212``); adjoint ``correlate(..., 'full')``, zero the overridden cells,
``np.add.at`` fold-back (``filter.py:214-220``).

Pinning (tests/test_oracle_filter.py): the reference's own known answers
(``tests/test_filter.py:13-213``: kernel placement in the interior, mirrored at edges, constant
edges), the adjoint identity <Hx, y> = <x, H^T y>, and golden filtered fields from the reference.
"""
from numbers import Number

import numpy as np
from scipy.signal import convolve, correlate


def filter_weights(radius, nel, esize=(1.0, 1.0, 1.0), relative_units=True):
    nx, ny, nz = nel
    dx, dy, dz = (1.0, 1.0, 1.0) if relative_units else esize
    ex = min(nx, int((radius - 1e-10 * dx) / dx))
    ey = min(ny, int((radius - 1e-10 * dy) / dy))
    ez = min(nz, int((radius - 1e-10 * dz) / dz))
    cx, cy, cz = np.meshgrid(np.arange(-ex, ex + 1) * dx, np.arange(-ey, ey + 1) * dy,
                             np.arange(-ez, ez + 1) * dz, indexing="ij")
    w = np.maximum(0.0, radius - np.sqrt(cx * cx + cy * cy + cz * cz))
    return w / np.sum(w)


class ConvFilter:
    """Padded-index convolution filter; ``bcs`` is ((xmin, xmax), (ymin, ymax), (zmin, zmax))."""

    def __init__(self, nel, weights, bcs=(("symmetric",) * 2,) * 3):
        self.nel = [int(v) for v in nel]
        self.weights = np.array(weights, dtype=float)
        while self.weights.ndim < 3:
            self.weights = self.weights[..., None]
        self.pad = [s // 2 for s in self.weights.shape]
        sx, sy, sz = [max(1, s) for s in self.nel]
        ei, ej, ek = np.meshgrid(np.arange(sx), np.arange(sy), np.arange(sz), indexing="ij")
        self.el3d = (ek * self.nel[1] + ej) * self.nel[0] + ei
        self.overrides = []
        arr = self.el3d
        for d in range(3):
            arr = self._pad_axis(arr, bcs[d][0], bcs[d][1], d)
        self.el3d_pad = arr

    def _pad_axis(self, idx, bc0, bc1, d):
        p = self.pad[d]
        padded_sizes = [n + 2 * q for n, q in zip(self.nel, self.pad)]
        if bc0 == "wrap" or bc1 == "wrap":
            pw = [(0, 0)] * 3
            pw[d] = (p if bc0 == "wrap" else 0, p if bc1 == "wrap" else 0)
            idx = np.pad(idx, pw, mode="wrap")
        for side, bc in ((1, bc1), (0, bc0)):
            pw = [(0, 0)] * 3
            pw[d] = (0, p) if side == 1 else (p, 0)
            if bc in ("edge", "symmetric"):
                idx = np.pad(idx, pw, mode=bc)
            elif isinstance(bc, Number) and not isinstance(bc, bool):
                idx = np.pad(idx, pw, mode="constant", constant_values=0)
                rng = [np.arange(max(1, s)) for s in padded_sizes]
                rng[d] = (p + self.nel[d] + np.arange(p)) if side == 1 else np.arange(p)
                ix = np.meshgrid(*rng, indexing="ij")
                if not all(np.asarray(i).size == 0 for i in ix):
                    self.overrides.append((tuple(ix), float(bc)))
        return idx

    def forward(self, x):
        xpad = x[self.el3d_pad]
        for ix, val in self.overrides:
            xpad[ix] = val
        y3d = convolve(xpad, self.weights, mode="valid")
        y = np.zeros_like(x)
        np.add.at(y, self.el3d, y3d)
        return y

    def adjoint(self, dy):
        g = correlate(dy[self.el3d], self.weights, mode="full")
        for ix, _ in self.overrides:
            g[ix] = 0
        dx = np.zeros_like(dy)
        np.add.at(dx, self.el3d_pad, g)
        return dx


def density_filter_matrix(nel, radius=2.0):
    """H of DensityFilter._calculate_h (pymoto/modules/filter.py:310-378): coo entries
    max(0, r - |i - j|) (element units) over the window |j - i|_inf <= int(r) clipped to the grid;
    nel = (nelx, nely, nelz or 0)."""
    import scipy.sparse as sps
    nx, ny, nz = nel[0], nel[1], max(nel[2], 1)
    d = int(radius)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    rows, cols, vals = [], [], []
    for dk in range(-d, d + 1):
        for dj in range(-d, d + 1):
            for di in range(-d, d + 1):
                ii, jj, kk = i + di, j + dj, k + dk
                ok = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny) & (kk >= 0) & (kk < nz)
                w = max(0.0, radius - np.sqrt(di * di + dj * dj + dk * dk))
                rows.append(((k * ny + j) * nx + i)[ok])
                cols.append(((kk * ny + jj) * nx + ii)[ok])
                vals.append(np.full(ok.sum(), w))
    n = nx * ny * nz
    return sps.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n)).tocsc()


class DensityFilterOracle:
    """y = H x / s, adjoint H (dy / s) (filter.py:242-267)."""

    def __init__(self, nel, radius=2.0, nonpadding=None):
        self.H = density_filter_matrix(nel, radius)
        self.Hs = np.asarray(self.H.sum(1)).ravel()
        if nonpadding is not None:
            inds = ~np.isin(np.arange(len(self.Hs)), nonpadding)
            self.Hs[inds] = np.max(self.Hs)

    def forward(self, x):
        return self.H @ x / self.Hs

    def adjoint(self, dy):
        return self.H @ (dy / self.Hs)
