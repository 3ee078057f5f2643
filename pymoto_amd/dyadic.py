"""Low-rank sensitivity carrier ``DyadicMatrix`` with device vectors.

Mirrors ``pymoto/common/dyadcarrier.py:34-525``: A = sum_k fac_k u_k (x) v_k.  The reference
copies ``fac*u`` on ``add_dyad``; here each term keeps a reference to its device vectors plus
the scalar factor, which the contraction kernel applies, so ``LinSolve`` can hand over
``DyadicMatrix(-lambda, u)`` without materialising ``-lambda`` (its solution tensors are never
written in place afterwards).  The per-element contraction with an element matrix
(``contract(K0, dofconn, dofconn)``, ``dyadcarrier.py:392-414``) is done by the assembly module's
HIP kernel; ``contract()`` without arguments is the trace sum_k fac_k u_k . v_k.
"""
import torch

from . import _lib


class DyadicMatrix:
    ndim = 2

    def __init__(self, u=None, v=None, shape=(-1, -1), fac=1.0):
        self._terms = []
        self.ulen, self.vlen = shape
        if u is not None:
            self.add_dyad(u, v, fac)

    @property
    def shape(self):
        return (self.ulen, self.vlen)

    @property
    def size(self):
        return 0 if self.ulen < 0 or self.vlen < 0 else self.ulen * self.vlen

    @property
    def n_dyads(self):
        return len(self._terms)

    @property
    def u(self):
        return [t[0] for t in self._terms]

    @property
    def v(self):
        return [t[1] for t in self._terms]

    def terms(self):
        """(u_k, v_k, fac_k) triples."""
        return list(self._terms)

    def add_dyad(self, u, v=None, fac=None):
        ul = u if isinstance(u, (list, tuple)) else [u]
        vl = ul if v is None else (v if isinstance(v, (list, tuple)) else [v])
        if len(ul) != len(vl):
            raise TypeError(f"Number of vectors in u ({len(ul)}) and v({len(vl)}) should be equal")
        for ui, vi in zip(ul, vl):
            ui, vi = _lib.as_device(ui), _lib.as_device(vi)
            if ui.ndim != 1 or vi.ndim != 1:
                raise NotImplementedError("device DyadicMatrix terms must be 1-D vectors")
            if self.ulen < 0:
                self.ulen = ui.shape[-1]
            if self.vlen < 0:
                self.vlen = vi.shape[-1]
            if ui.shape[-1] != self.ulen or vi.shape[-1] != self.vlen:
                raise TypeError("vector size not conforming to the dyad size")
            self._terms.append((ui.contiguous(), vi.contiguous(), 1.0 if fac is None else float(fac)))
        return self

    def __iadd__(self, other):
        for u, v, f in other.terms():
            self._terms.append((u, v, f))
        if self.ulen < 0:
            self.ulen, self.vlen = other.shape
        return self

    def add_sensitivity(self, other):
        return self.__iadd__(other)

    def __neg__(self):
        out = DyadicMatrix(shape=self.shape)
        out._terms = [(u, v, -f) for u, v, f in self._terms]
        return out

    def __mul__(self, s):
        out = DyadicMatrix(shape=self.shape)
        out._terms = [(u, v, f * float(s)) for u, v, f in self._terms]
        return out

    __rmul__ = __mul__

    def copy(self):
        out = DyadicMatrix(shape=self.shape)
        out._terms = list(self._terms)
        return out

    def __deepcopy__(self, memo):
        return self.copy()

    def contract(self, mat=None, rows=None, cols=None):
        if mat is not None or rows is not None or cols is not None:
            raise NotImplementedError("device DyadicMatrix contracts with the stiffness element matrix through "
                                      "AssembleStiffness; only the trace contraction is available here")
        lib = _lib.lib()
        out = torch.empty(1, dtype=torch.float64, device=_lib.device())
        total = 0.0
        for u, v, f in self._terms:
            _lib.check(lib.pmt_dot(u.numel(), _lib.ptr(u), _lib.ptr(v), _lib.ptr(out), _lib.ptr(_lib.reduce_work()),
                                   _lib.stream()), "dot")
            total += f * float(out.item())
        return total
