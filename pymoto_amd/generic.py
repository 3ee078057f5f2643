"""``MathExpression`` (material interpolation) and ``EinSum`` (compliance) on the device.

``MathExpression`` (``pymoto/modules/generic.py:14-140``) parses the expression with sympy exactly
as the reference does (``^`` -> ``**``, lower case, ``inp0`` for the first input) and
differentiates it symbolically.  The device path covers the power-law family a + b*inp0**p that
SIMP uses (``ex_compliance.py:89``: ``f"{xmin} + {1-xmin}*inp0^3"``) and the Heaviside projection
of the robust formulation (``ex_compliance_robust.py:139``, beta constant or a second input); the
constants are taken from the same printed form that ``sympy.lambdify`` hands to numpy, and one HIP
kernel evaluates each.  Other expressions raise ``NotImplementedError``.

``EinSum`` (``generic.py:143-226``) covers the dot product ``'i,i->'`` of the compliance c = u.f
with a deterministic device reduction; its adjoint returns dc*f and dc*u.
"""
import numpy as np
import torch

from . import _lib
from .core import Module


def _printed_float(v) -> float:
    """The double numpy sees for a sympy number in lambdify-generated code."""
    from sympy.printing.numpy import NumPyPrinter
    return float(eval(NumPyPrinter().doprint(v), {"numpy": np}))


def _parse_powlaw(expression: str, names=()):
    from sympy import Symbol, Wild
    from sympy.parsing.sympy_parser import parse_expr
    text = expression.replace("^", "**").lower()
    expr = parse_expr(text)
    x = Symbol("inp0")
    repl = {Symbol(n.lower()): x for n in names if n}
    expr = expr.subs(repl)
    if expr.free_symbols - {x}:
        raise NotImplementedError(f"device MathExpression supports one input; got {expr.free_symbols}")
    a, b, p = Wild("a", exclude=[x]), Wild("b", exclude=[x]), Wild("p", exclude=[x])
    m = expr.match(a + b * x ** p)
    if not m:
        raise NotImplementedError(f"device MathExpression supports a + b*inp0**p, not '{expression}'")
    dexpr = expr.diff(x)
    c, q = Wild("c", exclude=[x]), Wild("q", exclude=[x])
    md = dexpr.match(c * x ** q) if dexpr != 0 else {c: 0, q: 0}
    if md is None or not m[p].is_number or not md[q].is_number:
        raise NotImplementedError(f"device MathExpression supports a + b*inp0**p, not '{expression}'")
    coeffs = tuple(_printed_float(v) for v in (m[a], m[b], m[p], md[c], md[q]))
    # guard against a spurious symbolic match
    xs = np.array([0.13, 0.5, 0.77, 1.0])
    ref = np.array([float(expr.subs(x, v)) for v in xs])
    got = coeffs[0] + coeffs[1] * xs ** coeffs[2]
    if not np.allclose(ref, got, rtol=1e-12, atol=1e-14):
        raise NotImplementedError(f"device MathExpression could not reduce '{expression}' to a + b*inp0**p")
    return coeffs


def _heaviside(x, beta, eta):
    return (np.tanh(beta * eta) + np.tanh(beta * (x - eta))) / (np.tanh(beta * eta) + np.tanh(beta * (1 - eta)))


def _parse_heaviside(expression: str, names=()):
    """(beta or None, eta) if the expression is the Heaviside projection of the robust formulation
    (ex_compliance_robust.py:139) in inp0 with a constant beta or beta = inp1."""
    from sympy import Symbol, tanh
    from sympy.parsing.sympy_parser import parse_expr
    expr = parse_expr(expression.replace("^", "**").lower())
    x, bsym = Symbol("inp0"), Symbol("inp1")
    repl = {Symbol(n.lower()): s for n, s in zip(names, (x, bsym)) if n}
    expr = expr.subs(repl)
    if not expr.free_symbols or expr.free_symbols - {x, bsym}:
        return None
    args = [t.args[0] for t in expr.atoms(tanh)]
    with_x = [a for a in args if x in a.free_symbols]
    if not with_x:
        return None
    slope = with_x[0].diff(x)
    beta = None if bsym in expr.free_symbols else slope
    if beta is not None and not beta.is_number:
        return None
    b_eval = bsym if beta is None else beta
    eta = -(with_x[0] - slope * x) / b_eval
    eta = eta.simplify()
    if not eta.is_number:
        return None
    beta_v = None if beta is None else _printed_float(beta)
    eta_v = _printed_float(eta)
    for xv in (0.05, 0.37, 0.5, 0.93):
        for bv in ((1.5, 8.0) if beta is None else (beta_v,)):
            ref = float(expr.subs({x: xv, bsym: bv}))
            if not np.isclose(ref, _heaviside(xv, bv, eta_v), rtol=1e-12, atol=1e-14):
                return None
    return beta_v, eta_v


class MathExpression(Module):
    """y = f(inp0[, inp1]) for the families the density chain uses, derivative by sympy:

    * power law a + b*inp0**p (SIMP, ``ex_compliance.py:89``);
    * Heaviside projection (tanh(beta*eta) + tanh(beta*(inp0-eta))) / (tanh(beta*eta) + tanh(beta*(1-eta)))
      with a constant beta or beta = inp1 (``ex_compliance_robust.py:139``).
    """

    def __init__(self, expression: str):
        self.expression = expression
        self.coeffs = None
        self.heaviside = None

    def _parse(self, nargs):
        names = [getattr(s, "tag", "") for s in (self.sig_in or [])]
        hv = _parse_heaviside(self.expression, names)
        if hv is not None:
            if (hv[0] is None) != (nargs == 2):
                raise NotImplementedError("Heaviside MathExpression: beta must be inp1 (2 inputs) or a constant")
            self.heaviside = hv
            return
        if nargs != 1:
            raise NotImplementedError("device MathExpression takes one input except for the Heaviside projection")
        self.coeffs = _parse_powlaw(self.expression, names)

    def _beta(self):
        b = self.heaviside[0]
        return float(np.asarray(self.sig_in[1].state if b is None else b, dtype=np.float64).reshape(()))

    def __call__(self, *args):
        if len(args) not in (1, 2):
            raise NotImplementedError("device MathExpression takes one or two inputs")
        if self.coeffs is None and self.heaviside is None:
            self._parse(len(args))
        x = _lib.as_device(args[0])
        y = torch.empty_like(x)
        if self.heaviside is not None:
            beta = float(np.asarray(args[1] if self.heaviside[0] is None else self.heaviside[0],
                                    dtype=np.float64).reshape(()))
            _lib.check(_lib.lib().pmt_heaviside_fwd(x.numel(), beta, self.heaviside[1], _lib.ptr(x), _lib.ptr(y),
                                                    _lib.stream()), "heaviside")
            return y
        a, b, p, _, _ = self.coeffs
        _lib.check(_lib.lib().pmt_powlaw_fwd(x.numel(), a, b, p, _lib.ptr(x), _lib.ptr(y), _lib.stream()),
                   "power law")
        return y

    def _sensitivity(self, df_dy):
        x = _lib.as_device(self.sig_in[0].state)
        dy = _lib.as_device(df_dy)
        dx = torch.empty_like(x)
        as_np = isinstance(self.sig_in[0].state, np.ndarray)
        if self.heaviside is not None:
            beta, eta = self._beta(), self.heaviside[1]
            two = self.heaviside[0] is None
            db = torch.empty_like(x) if two else None
            lib = _lib.lib()
            _lib.check(lib.pmt_heaviside_bwd(x.numel(), beta, eta, _lib.ptr(x), _lib.ptr(dy), _lib.ptr(dx), 0,
                                             _lib.ptr(db), _lib.stream()), "heaviside adjoint")
            dxo = dx.cpu().numpy() if as_np else dx
            if not two:
                return dxo
            ones = torch.ones_like(db)
            out = torch.empty(1, dtype=torch.float64, device=db.device)
            _lib.check(lib.pmt_dot(db.numel(), _lib.ptr(db), _lib.ptr(ones), _lib.ptr(out),
                                   _lib.ptr(_lib.reduce_work()), _lib.stream()), "dot")
            return [dxo, float(out.item())]
        _, _, _, c, q = self.coeffs
        _lib.check(_lib.lib().pmt_powlaw_bwd(x.numel(), c, q, _lib.ptr(x), _lib.ptr(dy), _lib.ptr(dx), 0,
                                             _lib.stream()), "power law adjoint")
        return dx.cpu().numpy() if as_np else dx


class EinSum(Module):
    """The contractions of ``generic.py:143-226`` that the compliance workloads use, on the device:

    * ``'i,i->'``   dot product (compliance ``f.u``, ``ex_compliance.py``);
    * ``'ij,ij->j'`` column-wise dot products (compliance per load case, ``ex_compliant_mechanism.py:83``);
    * ``'ij,ij->'``  sum over all columns;
    * ``'i->'``      vector sum (volume, ``ex_compliance_robust.py``).

    Sensitivities follow the adjoint einsum of ``generic.py:213-226``: for a dot product each input
    receives the other times the output sensitivity (per column for ``'ij,ij->j'``), for a sum every
    entry receives the output sensitivity.  Other expressions raise ``NotImplementedError``."""

    _supported = {(("i", "i"), ""), (("ij", "ij"), "j"), (("ij", "ij"), ""), (("i",), "")}

    def __init__(self, expression: str):
        self.expr = expression
        cmd = expression.replace(" ", "").split("->")
        self.indices_in = cmd[0].split(",")
        self.indices_out = cmd[1] if "->" in expression else ""
        if (tuple(self.indices_in), self.indices_out) not in self._supported:
            raise NotImplementedError(f"device EinSum supports 'i,i->', 'ij,ij->j', 'ij,ij->' and 'i->', "
                                      f"not '{expression}'")

    @staticmethod
    def _dot(u, v):
        out = torch.empty(1, dtype=torch.float64, device=u.device)
        _lib.check(_lib.lib().pmt_dot(u.numel(), _lib.ptr(u), _lib.ptr(v), _lib.ptr(out),
                                      _lib.ptr(_lib.reduce_work()), _lib.stream()), "dot")
        return out

    @staticmethod
    def _scaled(v, a):
        d = torch.empty_like(v)
        _lib.check(_lib.lib().pmt_axpby(v.numel(), float(a), _lib.ptr(v), 0.0, _lib.ptr(d), _lib.stream()), "scale")
        return d

    def __call__(self, *args):
        ops = [_lib.as_device(a) for a in args]
        if len(ops) == 1:  # 'i->'
            u = ops[0]
            if u.ndim != 1:
                raise ValueError("EinSum 'i->' needs a vector")
            return np.float64(self._dot(u, torch.ones_like(u)).item())
        u, v = ops
        if u.shape != v.shape or u.ndim != len(self.indices_in[0]):
            raise ValueError(f"EinSum '{self.expr}' needs two operands of equal shape and rank "
                             f"{len(self.indices_in[0])}")
        if self.indices_out == "":
            return np.float64(self._dot(u, v).item())
        ut, vt = u.t().contiguous(), v.t().contiguous()  # column j contiguous
        return torch.cat([self._dot(ut[j], vt[j]) for j in range(ut.shape[0])])

    def _sensitivity(self, df_in):
        inps = self.get_input_states(as_list=True)
        if len(inps) == 1:
            u = _lib.as_device(inps[0])
            d = self._scaled(torch.ones_like(u), float(df_in))
            return d.cpu().numpy() if isinstance(inps[0], np.ndarray) else d
        a, b = inps
        res = []
        for other, mine in ((b, a), (a, b)):
            o = _lib.as_device(other)
            if self.indices_out == "j":
                dfj = np.asarray(df_in.cpu().numpy() if isinstance(df_in, torch.Tensor) else df_in,
                                 dtype=np.float64).reshape(-1)
                ot = o.t().contiguous()
                d = torch.stack([self._scaled(ot[j], dfj[j]) for j in range(ot.shape[0])], dim=1)
            else:
                d = self._scaled(o, float(df_in))
            res.append(d.cpu().numpy() if isinstance(mine, np.ndarray) else d)
        return res
