"""``FilterConv`` on the device (``pymoto/modules/filter.py:8-220``).

Same constructor and boundary options as the reference: ``radius`` (linear-hat weights
max(0, r - |d|) on a box of half-width floor((r - 1e-10 h)/h) clipped to the domain, normalised to
sum 1, ``filter.py:189-205``) or explicit odd-sized ``weights``; per face 'symmetric' (default),
'edge', 'wrap' or a constant value.  Setup turns the reference's axis-by-axis ``np.pad`` of the
element-index array into one small index map per axis (padding along an axis applies the same
1-D map to every line) plus constant cells, and the inverse lists needed by the adjoint; the
forward and adjoint run in HIP kernels.
"""
from numbers import Number

import numpy as np
import torch

from . import _lib
from .core import Module
from .domain import VoxelDomain


def _axis_map(n: int, p: int, bc0, bc1):
    """Padded index map of one axis: np.pad semantics (wrap first, then the max side, then the min side)."""
    idx = np.arange(max(n, 1), dtype=np.int64)
    val = np.zeros(idx.size)
    if p == 0:
        return idx, val
    if p > max(n, 1):
        raise NotImplementedError("filter half-width larger than the domain extent is not supported")
    if bc0 == "wrap" or bc1 == "wrap":
        pw = (p if bc0 == "wrap" else 0, p if bc1 == "wrap" else 0)
        idx = np.pad(idx, pw, mode="wrap")
        val = np.pad(val, pw, mode="constant")
    for side, bc in ((1, bc1), (0, bc0)):
        pw = (0, p) if side == 1 else (p, 0)
        if bc in ("edge", "symmetric"):
            idx = np.pad(idx, pw, mode=bc)
            val = np.pad(val, pw, mode=bc)
        elif isinstance(bc, Number) and not isinstance(bc, bool):
            idx = np.pad(idx, pw, mode="constant", constant_values=-1)
            val = np.pad(val, pw, mode="constant", constant_values=float(bc))
        elif bc != "wrap":
            raise ValueError(f"unknown boundary condition {bc!r}")
    return idx, val


def _inverse_lists(amap: np.ndarray, n: int):
    ptr = np.zeros(n + 1, dtype=np.int64)
    valid = amap >= 0
    np.add.at(ptr, amap[valid] + 1, 1)
    ptr = np.cumsum(ptr)
    order = np.argsort(amap[valid], kind="stable")
    idx = np.flatnonzero(valid)[order]
    return ptr, idx


class FilterConv(Module):
    """Density filter based on convolution (``filter.py:8-220``)."""

    def __init__(self, domain: VoxelDomain, radius: float = None, relative_units: bool = True, weights=None,
                 xmin_bc="symmetric", xmax_bc="symmetric", ymin_bc="symmetric", ymax_bc="symmetric",
                 zmin_bc="symmetric", zmax_bc="symmetric"):
        self.domain = domain
        if (weights is None) == (radius is None):
            raise ValueError("Only one of arguments 'filter_radius' or 'weights' must be provided.")
        if weights is not None:
            w = np.array(weights, dtype=np.float64)
            while w.ndim < 3:
                w = w[..., None]
            for s in w.shape:
                assert s % 2 == 1, "Size of weights must be uneven"
            self.weights = w
        else:
            self.set_filter_radius(radius, relative_units)
        self.bcs = ((xmin_bc, xmax_bc), (ymin_bc, ymax_bc), (zmin_bc, zmax_bc))
        self._setup()

    def set_filter_radius(self, radius: float, relative_units: bool = True):
        dx, dy, dz = (1.0, 1.0, 1.0) if relative_units else self.domain.element_size
        nx, ny, nz = self.domain.nelx, self.domain.nely, self.domain.nelz
        ex = min(nx, int((radius - 1e-10 * dx) / dx))
        ey = min(ny, int((radius - 1e-10 * dy) / dy))
        ez = min(nz, int((radius - 1e-10 * dz) / dz))
        cx, cy, cz = np.meshgrid(np.arange(-ex, ex + 1) * dx, np.arange(-ey, ey + 1) * dy,
                                 np.arange(-ez, ez + 1) * dz, indexing="ij")
        self.weights = np.maximum(0.0, radius - np.sqrt(cx * cx + cy * cy + cz * cz))
        self.weights /= np.sum(self.weights)
        if hasattr(self, "bcs"):
            self._setup()

    def _setup(self):
        d = self.domain
        self.nel3 = (d.nelx, d.nely, d.nelz if d.dim == 3 else 0)
        sizes = (d.nelx, d.nely, max(d.nelz, 1) if d.dim == 3 else 1)
        if d.dim != 3 and self.weights.shape[2] != 1:
            raise ValueError("a 2D domain needs weights of size 1 in z")
        maps, vals, ptrs, idxs = [], [], [], []
        offset = 0
        for ax in range(3):
            p = self.weights.shape[ax] // 2
            m, v = _axis_map(sizes[ax], p, *self.bcs[ax])
            ptr, idx = _inverse_lists(m, sizes[ax])
            # the tiled adjoint kernel takes every element farther than the half width p from both faces
            # to have only its own padded cell (p + a); the padding modes put the others within p of a face
            inner = np.arange(p, sizes[ax] - p)
            assert np.all(np.diff(ptr)[inner] == 1) and np.all(idx[ptr[inner]] == inner + p)
            maps.append(m)
            vals.append(v)
            ptrs.append(ptr)
            idxs.append(idx)
        inv_idx = []
        inv_ptr = []
        for ax in range(3):
            inv_ptr.append(ptrs[ax] + offset)
            inv_idx.append(idxs[ax])
            offset += idxs[ax].size
        self._map = np.concatenate(maps).astype(np.int32)
        self._cval = np.concatenate(vals)
        self._inv_ptr = np.concatenate(inv_ptr).astype(np.int32)
        self._inv_idx = np.concatenate(inv_idx).astype(np.int32)
        self._dev = None

    def _device_data(self):
        dev = torch.cuda.current_device()
        if self._dev is None or self._dev[0] != dev:
            self._dev = (dev, _lib.as_device(self.weights.ravel()), _lib.as_device(self._map, torch.int32),
                         _lib.as_device(self._cval), _lib.as_device(self._inv_ptr, torch.int32),
                         _lib.as_device(self._inv_idx, torch.int32))
        return self._dev[1:]

    def __call__(self, x):
        lib = _lib.lib()
        xd = _lib.as_device(x).reshape(-1)
        if xd.numel() != self.domain.nel:
            raise ValueError(f"input of size {xd.numel()} does not match the {self.domain.nel} elements")
        w, m, cv, _, _ = self._device_data()
        y = torch.empty_like(xd)
        _lib.check(lib.pmt_filter_fwd(*self.nel3, *self.weights.shape, _lib.ptr(w), _lib.ptr(m), _lib.ptr(cv),
                                      _lib.ptr(xd), _lib.ptr(y), _lib.stream()), "filter forward")
        return y

    def _sensitivity(self, dfdv):
        lib = _lib.lib()
        dy = _lib.as_device(dfdv).reshape(-1)
        w, _, _, ip, ii = self._device_data()
        dx = torch.empty_like(dy)
        _lib.check(lib.pmt_filter_bwd(*self.nel3, *self.weights.shape, _lib.ptr(w), _lib.ptr(ip), _lib.ptr(ii),
                                      _lib.ptr(dy), _lib.ptr(dx), _lib.stream()), "filter adjoint")
        xs = self.sig_in[0].state
        return dx.cpu().numpy() if isinstance(xs, np.ndarray) else dx


class DensityFilter(Module):
    """Row-normalised density filter y = H x / s (``filter.py:223-378``), H_ij = max(0, r - |i - j|)
    in element units over the window |j - i|_inf <= int(r) clipped to the domain, s = H 1; elements
    not listed in ``nonpadding`` get s = max(s) (mimics zero padding outside the domain).  The
    adjoint is H (dy / s) (H is symmetric)."""

    def __init__(self, domain: VoxelDomain, radius: float = 2.0, nonpadding=None):
        self.domain = domain
        self.radius = float(radius)
        self.nonpadding = nonpadding
        self._s_inv = None

    def _grid(self):
        d = self.domain
        return d.nelx, d.nely, d.nelz if d.dim == 3 else 0

    def _apply(self, x, si=None, so=None):
        y = torch.empty_like(x)
        _lib.check(_lib.lib().pmt_density_filter(*self._grid(), self.radius, _lib.ptr(x), _lib.ptr(si), _lib.ptr(so),
                                                 _lib.ptr(y), _lib.stream()), "density filter")
        return y

    def _scale(self, like):
        if self._s_inv is None:
            s = self._apply(torch.ones_like(like))
            if self.nonpadding is not None:
                keep = torch.zeros(s.numel(), dtype=torch.bool, device=s.device)
                keep[torch.as_tensor(np.asarray(self.nonpadding), device=s.device, dtype=torch.long)] = True
                s = torch.where(keep, s, s.max())
            self._s_inv = (1.0 / s).contiguous()
        return self._s_inv

    def __call__(self, x):
        xd = _lib.as_device(x)
        return self._apply(xd, so=self._scale(xd))

    def _sensitivity(self, dfdy):
        dy = _lib.as_device(dfdy)
        dx = self._apply(dy, si=self._scale(dy))
        return dx.cpu().numpy() if isinstance(self.sig_in[0].state, np.ndarray) else dx
