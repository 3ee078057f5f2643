"""Dataflow graph of the hot path: ``Signal``, ``SignalSlice``, ``Module``, ``Network``.

The contract is pyMOTO's Module API (``pymoto/core_objects.py:71-998``); the mechanics here are
this package's own:

* A module class writes its response as ``__call__(self, *inputs)`` and its adjoint as
  ``_sensitivity(self, *d_outputs)``, positional parameters only.  The metaclass ``_ModuleMeta``
  moves the class's ``__call__`` to ``_evaluate`` when the class is created (after checking the
  signature), and the single ``Module.__call__`` below wires signals around it: calling a module
  with at least one signal (or with nothing: a source) records the inputs, evaluates once, creates
  ``"{Class}_output{i}"`` signals and appends the module to every open ``Network``; calling it with
  plain values only just evaluates.
* A module object is wired once.  Wired modules are kept in a weak set (identity, not equality), so
  a second call with signals raises ``RuntimeError`` while a deep copy is a new, unwired module.
* ``sensitivity()`` is skipped when every output sensitivity is ``None``; results are accumulated
  into the input signals by ``add_sensitivity``, which stores a private copy of the first term so
  that two inputs never share one array.
* Anything with ``state/sensitivity/add_sensitivity/reset`` is a signal and anything with
  ``response/sensitivity/reset`` is a module, so reference pyMOTO objects mix with these.

Device tensors (``torch`` on the GPU) pass through unchanged; nothing here touches a device.
"""
import copy
import sys
import time
import warnings
import weakref
from abc import ABCMeta, abstractmethod
from inspect import Parameter, signature
from typing import Any, List, Optional

import numpy as np

__all__ = ["Signal", "SignalSlice", "Module", "Network", "make_signals"]


# --------------------------------------------------------------------------------------------
# helpers

def _as_list(value: Any) -> list:
    """``None`` -> [], list -> itself, tuple/set -> list, anything else -> [value]."""
    if value is None:
        return []
    if isinstance(value, list):
        return value
    if isinstance(value, (tuple, set)):
        return list(value)
    return [value]


def _is_signal(obj: Any) -> bool:
    return isinstance(obj, Signal) or all(hasattr(obj, a) for a in ("state", "sensitivity", "add_sensitivity", "reset"))


def _is_module(obj: Any) -> bool:
    return isinstance(obj, Module) or all(hasattr(obj, a) for a in ("response", "sensitivity", "reset"))


def _root(sig: Any) -> Any:
    """The signal that owns the data of ``sig`` (a slice points at its base)."""
    while isinstance(sig, SignalSlice):
        sig = sig.base
    return sig


def _signal_set(sigs: Any) -> set:
    return {_root(s) for s in _as_list(sigs) if _is_signal(s)}


def _where() -> str:
    """File and line of the first caller outside this file (for error messages)."""
    frame = sys._getframe(1)
    while frame is not None and frame.f_code.co_filename == __file__:
        frame = frame.f_back
    return "unknown location" if frame is None else f"{frame.f_code.co_filename}:{frame.f_lineno}"


def _rethrow(exc: BaseException, context: str):
    """Raise an exception of the same type with ``context`` appended to its message."""
    try:
        new = type(exc)(f"{exc}{context}")
    except Exception:  # exception types whose constructor takes other arguments
        raise exc
    raise new.with_traceback(exc.__traceback__) from None


def _arity(fn) -> tuple:
    """(minimum, maximum or None) positional arguments of ``fn``; keyword-only parameters and
    ``**kwargs`` are refused (``core_objects.py:408-448``)."""
    lo, hi = 0, 0
    for p in signature(fn).parameters.values():
        if p.kind in (Parameter.POSITIONAL_ONLY, Parameter.POSITIONAL_OR_KEYWORD):
            hi = None if hi is None else hi + 1
            lo += p.default is Parameter.empty
        elif p.kind == Parameter.VAR_POSITIONAL:
            hi = None
        else:
            raise SyntaxError(f"{getattr(fn, '__qualname__', fn)}: only positional parameters are allowed, "
                              f"'{p}' is not")
    return lo, hi


def _check_count(fn, n: int, what: str):
    lo, hi = _arity(fn)
    if n < lo or (hi is not None and n > hi):
        want = f"{lo}" if hi == lo else f"{lo}..{'any' if hi is None else hi}"
        raise TypeError(f"{getattr(fn, '__qualname__', fn)} takes {want} {what}, {n} would be passed")


def _zeros_like(state: Any) -> Any:
    if state is None:
        raise TypeError("a signal without a state has no shape to hold a sensitivity")
    if hasattr(state, "new_zeros"):  # torch tensor
        return state.new_zeros(state.shape)
    try:
        return np.zeros_like(state) if isinstance(state, np.ndarray) else state * 0
    except TypeError:
        raise TypeError(f"cannot create a zero sensitivity for a state of type {type(state).__name__}") from None


# --------------------------------------------------------------------------------------------
# signals

class Signal:
    """Carries a ``state`` and its ``sensitivity`` between modules (``core_objects.py:71-195``).

    ``keep_alloc`` (default: whether a sensitivity was given at construction) makes ``reset()``
    zero the sensitivity in place instead of dropping it.
    """

    def __init__(self, tag: str = "", state: Any = None, sensitivity: Any = None, min: Any = None, max: Any = None):
        self.tag = tag
        self.state = state
        self.sensitivity = sensitivity
        self.min = min
        self.max = max
        self.keep_alloc = sensitivity is not None
        self._origin = _where()

    def add_sensitivity(self, ds: Any):
        """Accumulate ``ds``: a private copy of the first term, in-place addition afterwards."""
        if ds is None:
            return self
        try:
            if self.sensitivity is None:
                self.sensitivity = copy.deepcopy(ds)
            elif hasattr(self.sensitivity, "add_sensitivity"):  # e.g. a dyadic carrier
                self.sensitivity.add_sensitivity(ds)
            else:
                self.sensitivity += ds
        except (TypeError, ValueError) as e:
            _rethrow(e, f"\n  while adding a {type(ds).__name__} to the sensitivity "
                        f"({type(self.sensitivity).__name__}) of signal '{self.tag}' created at {self._origin}")
        return self

    def reset(self, keep_alloc: Optional[bool] = None):
        """Zero the sensitivity in place (``keep_alloc``) or drop it."""
        if self.sensitivity is None:
            return self
        if not (self.keep_alloc if keep_alloc is None else keep_alloc):
            self.sensitivity = None
            return self
        try:
            self.sensitivity[...] = 0  # arrays and tensors keep their memory
            return self
        except (TypeError, AttributeError, IndexError):
            pass
        try:
            value = self.sensitivity
            value *= 0  # numbers (rebinds) and objects with in-place scaling
            self.sensitivity = value
            return self
        except TypeError:
            pass
        warnings.warn(f"signal '{self.tag}': a sensitivity of type {type(self.sensitivity).__name__} cannot be "
                      "zeroed in place; it is dropped instead", RuntimeWarning)
        self.sensitivity = None
        return self

    def __getitem__(self, index):
        return SignalSlice(self, index)

    def __repr__(self):
        what = "no state" if self.state is None else f"a {type(self.state).__name__} state"
        return f"Signal '{self.tag}' with {what}"


class SignalSlice(Signal):
    """View of part of a signal: reads and writes go to the base signal's arrays
    (``core_objects.py:198-315``)."""

    def __init__(self, base: Signal, index, tag: Optional[str] = None):
        self.base = base
        self.slice = index
        self.keep_alloc = False
        self.tag = f"{base.tag}[{index}]" if tag is None else tag
        self.min = self.max = None
        self._origin = _where()

    @property
    def state(self):
        return None if self.base.state is None else self.base.state[self.slice]

    @state.setter
    def state(self, value):
        self.base.state[self.slice] = value

    @property
    def sensitivity(self):
        return None if self.base.sensitivity is None else self.base.sensitivity[self.slice]

    @sensitivity.setter
    def sensitivity(self, value):
        if self.base.sensitivity is None:
            if value is None:
                return
            self.base.sensitivity = _zeros_like(self.base.state)
        self.base.sensitivity[self.slice] = 0 if value is None else value

    def add_sensitivity(self, ds: Any):
        if ds is None:
            return self
        if self.base.sensitivity is None:
            self.base.sensitivity = _zeros_like(self.base.state)
        self.base.sensitivity[self.slice] += ds
        return self

    def reset(self, keep_alloc: Optional[bool] = None):
        """Zero this part of the base sensitivity (the base keeps its array)."""
        if self.base.sensitivity is not None:
            self.base.sensitivity[self.slice] = 0
        return self


def make_signals(*tags: str, **tagged_states: Any):
    """``make_signals('a', 'b')`` -> two signals; ``make_signals(a=1.0)`` -> {tag: signal with state};
    both forms together -> (tuple, dict) (``core_objects.py:318-357``)."""
    plain = tuple(Signal(t) for t in tags)
    named = {t: Signal(t, v) for t, v in tagged_states.items()}
    if tags and tagged_states:
        return plain, named
    if tagged_states:
        return named
    return plain[0] if len(plain) == 1 else plain


# --------------------------------------------------------------------------------------------
# modules

_WIRED = weakref.WeakSet()  # module objects that have been wired to signals (identity)


class _ModuleMeta(ABCMeta):
    """Moves a module class's ``__call__`` (its response) to ``_evaluate`` and checks that the
    response and ``_sensitivity`` take positional parameters only."""

    def __new__(mcs, name, bases, ns, **kw):
        if any(isinstance(b, _ModuleMeta) for b in bases):  # not the root class Module itself
            response = ns.pop("__call__", None)
            if response is not None:
                _arity(response)
                ns["_evaluate"] = response
        if "_sensitivity" in ns:
            _arity(ns["_sensitivity"])
        return super().__new__(mcs, name, bases, ns, **kw)


class Module(metaclass=_ModuleMeta):
    """A node of the graph.  Subclasses define ``__call__`` (response from input states) and
    optionally ``_sensitivity`` (input sensitivities from output sensitivities) and ``_reset``.
    Subclasses need not call ``Module.__init__``."""

    sig_in: Optional[List] = None
    sig_out: Optional[List] = None

    @abstractmethod
    def _evaluate(self, *states):
        """The subclass's ``__call__``."""

    # -- wiring ------------------------------------------------------------------------------
    def __call__(self, *args):
        if args and not any(_is_signal(a) for a in args):
            return self._evaluate(*args)  # plain values: a function call, no wiring
        if self in _WIRED:
            raise RuntimeError(f"{type(self).__name__} at {hex(id(self))} is already connected; "
                               "create another module (or a deep copy) for other signals")
        self.connect(list(args))
        for net in Network._open:
            net.append(self)
        if not self.sig_out:
            return None
        return self.sig_out[0] if len(self.sig_out) == 1 else tuple(self.sig_out)

    def connect(self, sig_in, sig_out=None):
        """Wire ``sig_in`` (and ``sig_out``, default new signals), evaluate once, without adding the
        module to an open network.  Returns the module."""
        self._origin = _where()
        self.sig_in = _as_list(sig_in)
        values = _as_list(self._evaluate(*self._input_states()))
        if sig_out is None:
            self.sig_out = [Signal(f"{type(self).__name__}_output{i}") for i in range(len(values))]
        else:
            self.sig_out = _as_list(sig_out)
        _check_count(self._sensitivity, len(self.sig_out), "output sensitivities")
        self._store(values)
        _WIRED.add(self)
        return self

    def _input_states(self) -> list:
        return [s.state if _is_signal(s) else s for s in self.sig_in]

    def _store(self, values: list):
        if len(values) != len(self.sig_out):
            raise TypeError(f"{type(self).__name__} returned {len(values)} values for {len(self.sig_out)} "
                            "output signals")
        for s, v in zip(self.sig_out, values):
            s.state = v

    @property
    def n_in(self):
        return None if self.sig_in is None else len(self.sig_in)

    @property
    def n_out(self):
        return None if self.sig_out is None else len(self.sig_out)

    def _describe(self) -> str:
        def tags(sigs):
            if sigs is None:
                return "(not connected)"
            return ", ".join(str(getattr(s, "tag", s)) for s in sigs)
        return (f"\n  in module {type(self).__name__}({tags(self.sig_in)}) -> {tags(self.sig_out)}, "
                f"connected at {getattr(self, '_origin', 'unknown location')}")

    # -- evaluation --------------------------------------------------------------------------
    def response(self):
        """Re-evaluate the outputs from the current input states."""
        try:
            self._store(_as_list(self._evaluate(*self._input_states())))
        except Exception as e:
            _rethrow(e, self._describe())
        return self

    def sensitivity(self):
        """Add this module's contribution to the input sensitivities."""
        if not self.sig_in or not any(_is_signal(s) for s in self.sig_in):
            return self  # nothing to propagate into
        d_out = [getattr(s, "sensitivity", None) for s in self.sig_out or []]
        if d_out and all(d is None for d in d_out):
            return self
        try:
            d_in = self._sensitivity(*d_out)
            d_in = list(d_in) if isinstance(d_in, (list, tuple)) else [d_in]
            if len(d_in) != len(self.sig_in):
                raise TypeError(f"_sensitivity returned {len(d_in)} values for {len(self.sig_in)} inputs")
            for s, d in zip(self.sig_in, d_in):
                if _is_signal(s):
                    s.add_sensitivity(d)
        except Exception as e:
            _rethrow(e, self._describe())
        return self

    def reset(self):
        """Clear the sensitivities of the connected signals and the module's own state."""
        for s in (self.sig_out or []) + (self.sig_in or []):
            if _is_signal(s):
                s.reset()
        self._reset()
        return self

    def get_input_states(self, as_list: bool = False):
        if self.sig_in is None:
            return None
        states = self._input_states()
        return states[0] if len(states) == 1 and not as_list else states

    def get_output_states(self, as_list: bool = False):
        if self.sig_out is None:
            return None
        states = [s.state for s in self.sig_out]
        return states[0] if len(states) == 1 and not as_list else states

    def get_output_sensitivities(self, as_list: bool = False):
        if self.sig_out is None:
            return None
        sens = [s.sensitivity for s in self.sig_out]
        return sens[0] if len(sens) == 1 and not as_list else sens

    def _sensitivity(self, *d_outputs):
        warnings.warn(f"{type(self).__name__} defines no _sensitivity; its inputs receive none", RuntimeWarning)
        return [None] * len(self.sig_in)

    def _reset(self):
        pass

    def __repr__(self):
        return f"{type(self).__name__} module at {hex(id(self))}"


# --------------------------------------------------------------------------------------------
# networks

class Network(Module):
    """An ordered list of modules evaluated as one: responses in order, sensitivities in reverse
    (``core_objects.py:750-994``).  Inside ``with Network() as fn:`` every module called with signals
    is appended to ``fn`` (and to any enclosing open network)."""

    _open: List["Network"] = []

    def __init__(self, *modules, print_timing=False):
        self.mods = []
        self.print_timing = print_timing
        self._origin = _where()
        self.append(*modules)

    def _evaluate(self, *args):
        raise RuntimeError("a Network is built from modules; it cannot be called with signals")

    def connect(self, sig_in, sig_out=None):
        self._evaluate()

    def __enter__(self):
        if any(n is self for n in Network._open):
            raise ValueError("this Network is already open")
        Network._open.append(self)
        return self

    def __exit__(self, *exc):
        Network._open.remove(self)
        return False

    def append(self, *modules):
        """Append modules; returns the output signal(s) of the last one."""
        mods = [m for group in modules for m in _as_list(group)]
        for i, m in enumerate(mods):
            if not _is_module(m):
                raise TypeError(f"argument {i} of append ({type(m).__name__}) is not a module")
        if not mods:
            return None
        self.mods.extend(mods)
        last = mods[-1].sig_out or []
        return last[0] if len(last) == 1 else last

    def _timed(self, calls, label):
        if self.print_timing is False:
            for call, _ in calls:
                call()
            return
        floor = 0.0 if self.print_timing is True else float(self.print_timing)
        start = time.perf_counter()
        for call, name in calls:
            t = time.perf_counter()
            call()
            dt = time.perf_counter() - t
            if dt >= floor:
                print(f"-- {label} of {name}: {dt:.6f} s")
        total = time.perf_counter() - start
        if total >= floor:
            print(f"-- {label} total: {total:.6f} s")

    def response(self):
        self._timed([(m.response, type(m).__name__) for m in self.mods], "response")
        return self

    def sensitivity(self):
        self._timed([(m.sensitivity, type(m).__name__) for m in reversed(self.mods)], "sensitivity")
        return self

    def reset(self):
        for m in reversed(self.mods):
            m.reset()
        return self

    def __len__(self):
        return len(self.mods)

    def __iter__(self):
        return iter(self.mods)

    @staticmethod
    def _unwrap(net: "Network"):
        if len(net) == 0:
            return None
        return net.mods[0] if len(net) == 1 else net

    def __getitem__(self, sel):
        """An int gives a module; a slice of positions gives a sub-network (or its only module, or
        None); a slice of signals ``fn[a:b]`` gives the modules between those signals."""
        if isinstance(sel, slice) and (_is_signal(sel.start) or _is_signal(sel.stop)):
            if sel.step is not None:
                raise IndexError("a slice of signals takes no step")
            for end in (sel.start, sel.stop):
                if end is not None and not _is_signal(end):
                    raise IndexError(f"a slice of signals cannot end at a {type(end).__name__}")
            return self._unwrap(self.get_input_cone(sel.start).get_output_cone(sel.stop))
        if isinstance(sel, slice):
            return self._unwrap(Network(self.mods[sel]))
        return self.mods[sel]

    def _flows(self):
        consumed, produced = set(), set()
        for m in self.mods:
            consumed |= _signal_set(m.sig_in)
            produced |= _signal_set(m.sig_out)
        return consumed, produced

    @property
    def sig_in(self):
        """Signals read by the network but produced by none of its modules."""
        consumed, produced = self._flows()
        return list(consumed - produced)

    @property
    def sig_out(self):
        """Signals produced by the network but read by none of its modules."""
        consumed, produced = self._flows()
        return list(produced - consumed)

    def get_input_cone(self, fromsig=None, frommod=None):
        """Modules that depend on ``fromsig`` or on the modules ``frommod`` (those included), in order;
        the whole network when both are empty."""
        reached = _signal_set(fromsig)
        starts = set(_as_list(frommod))
        if not reached and not starts:
            return self
        cone = Network()
        for m in self.mods:
            if m in starts or reached & _signal_set(m.sig_in):
                reached |= _signal_set(m.sig_out)
                cone.mods.append(m)
        return cone

    def get_output_cone(self, tosig=None, tomod=None):
        """Modules that ``tosig`` or the modules ``tomod`` (those included) depend on, in order; the
        whole network when both are empty."""
        needed = _signal_set(tosig)
        ends = set(_as_list(tomod))
        if not needed and not ends:
            return self
        picked = []
        for m in reversed(self.mods):
            if m in ends or needed & _signal_set(m.sig_out):
                needed |= _signal_set(m.sig_in)
                picked.append(m)
        return Network(picked[::-1])

    def get_subset(self, fromsig=None, tosig=None, include_sinks: bool = True, include_sources: bool = True):
        """The modules on the paths from ``fromsig`` to ``tosig`` (``core_objects.py:947-998``); with
        ``include_sources`` also the source modules (no inputs) whose outputs reach ``tosig`` and the
        modules in between, with ``include_sinks`` the sink modules (no outputs) that ``fromsig``
        reaches.  An empty end defaults to the open ends of the path between them."""
        starts, ends = _signal_set(fromsig), _signal_set(tosig)
        if not starts and not ends and include_sinks and include_sources:
            return self
        core = self.get_input_cone(starts).get_output_cone(ends)
        if len(core) == 0:
            return Network()
        starts = starts or set(core.sig_in)
        ends = ends or set(core.sig_out)
        sources, sinks = [], []
        if include_sources:
            # a source joins (with the modules between it and the path) when it reaches the end signals
            sources = [m for m in self.mods
                       if not m.sig_in and len(self.get_input_cone(frommod=m).get_output_cone(ends)) > 0]
        if include_sinks:
            # a sink joins (likewise) when the start signals reach it
            sinks = [m for m in self.mods
                     if not m.sig_out and len(self.get_output_cone(tomod=m).get_input_cone(starts)) > 0]
        return self.get_input_cone(starts, sources).get_output_cone(ends, sinks)

    def __repr__(self):
        return f"Network of {len(self.mods)} modules at {hex(id(self))}"
