"""``MathExpression`` (material interpolation) and ``EinSum`` (compliance) on the device.

``MathExpression`` (``pymoto/modules/generic.py:14-140``) parses the expression with sympy exactly
as the reference does (``^`` -> ``**``, lower case, ``inp0`` for the first input) and
differentiates it symbolically.  The device path covers the power-law family a + b*inp0**p that
SIMP uses (``ex_compliance.py:89``: ``f"{xmin} + {1-xmin}*inp0^3"``); its value and derivative
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


class MathExpression(Module):
    """y = f(inp0) for the power-law family a + b*inp0**p (SIMP); derivative by sympy."""

    def __init__(self, expression: str):
        self.expression = expression
        self.coeffs = None

    def __call__(self, *args):
        if len(args) != 1:
            raise NotImplementedError("device MathExpression takes exactly one input")
        if self.coeffs is None:
            names = [getattr(s, "tag", "") for s in (self.sig_in or [])]
            self.coeffs = _parse_powlaw(self.expression, names)
        a, b, p, _, _ = self.coeffs
        x = _lib.as_device(args[0])
        y = torch.empty_like(x)
        _lib.check(_lib.lib().pmt_powlaw_fwd(x.numel(), a, b, p, _lib.ptr(x), _lib.ptr(y), _lib.stream()),
                   "power law")
        return y

    def _sensitivity(self, df_dy):
        _, _, _, c, q = self.coeffs
        x = _lib.as_device(self.sig_in[0].state)
        dy = _lib.as_device(df_dy)
        dx = torch.empty_like(x)
        _lib.check(_lib.lib().pmt_powlaw_bwd(x.numel(), c, q, _lib.ptr(x), _lib.ptr(dy), _lib.ptr(dx), 0,
                                             _lib.stream()), "power law adjoint")
        return dx.cpu().numpy() if isinstance(self.sig_in[0].state, np.ndarray) else dx


class EinSum(Module):
    """Dot product ``'i,i->'`` on the device (``generic.py:143-226``)."""

    def __init__(self, expression: str):
        self.expr = expression
        cmd = expression.replace(" ", "").split("->")
        self.indices_in = cmd[0].split(",")
        self.indices_out = cmd[1] if "->" in expression else ""
        if self.indices_in != ["i", "i"] or self.indices_out != "":
            raise NotImplementedError(f"device EinSum supports 'i,i->' only, not '{expression}'")

    def __call__(self, a, b):
        u, v = _lib.as_device(a), _lib.as_device(b)
        if u.shape != v.shape or u.ndim != 1:
            raise ValueError("EinSum 'i,i->' needs two vectors of equal length")
        out = torch.empty(1, dtype=torch.float64, device=u.device)
        _lib.check(_lib.lib().pmt_dot(u.numel(), _lib.ptr(u), _lib.ptr(v), _lib.ptr(out),
                                      _lib.ptr(_lib.reduce_work()), _lib.stream()), "dot")
        return np.float64(out.item())

    def _sensitivity(self, df_in):
        a, b = self.get_input_states(as_list=True)
        lib = _lib.lib()
        res = []
        for other, mine in ((b, a), (a, b)):
            o = _lib.as_device(other)
            d = torch.empty_like(o)
            _lib.check(lib.pmt_axpby(o.numel(), float(df_in), _lib.ptr(o), 0.0, _lib.ptr(d), _lib.stream()), "scale")
            res.append(d.cpu().numpy() if isinstance(mine, np.ndarray) else d)
        return res
