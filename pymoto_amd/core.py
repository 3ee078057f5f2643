"""Dataflow graph: Signal, Module, Network.

Host-side mirror of pyMOTO's core contract (``pymoto/core_objects.py:71-998``) so that
the device modules of this package can be composed exactly like pyMOTO modules:

* ``Signal`` carries ``state`` and ``sensitivity``; sensitivities accumulate through
  ``add_sensitivity`` (deep copy on first use, ``+=`` afterwards), ``reset`` zeroes or
  clears them (``core_objects.py:107-165``).
* ``Module.__call__(*signals)`` connects the module, evaluates the response once, creates
  ``"{Class}_output{i}"`` signals and appends the module to every active ``Network``
  (``core_objects.py:481-548``). ``response()`` re-evaluates, ``sensitivity()`` back-propagates
  and is skipped when every output sensitivity is ``None`` (``core_objects.py:625-674``).
* ``Network`` runs responses in order and sensitivities in reverse (``core_objects.py:803-829``).

Device data (``torch`` tensors on the GPU) lives in ``Signal.state`` unchanged; nothing in this
file touches a device.  Duck typing is kept: any object with ``state/sensitivity/add_sensitivity/reset``
is a signal and any object with ``response/sensitivity/reset`` is a module, so modules of this
package can be appended to a pyMOTO ``Network`` and vice versa (``core_objects.py:360-381``).
"""
import copy
import inspect
import sys
import time
import warnings
from abc import ABC, abstractmethod
from typing import Any, List


def _parse_to_list(*args: Any) -> list:
    """Normalise ``None``/scalar/tuple/set/list arguments to a list (``pymoto/utils.py:5-20``)."""
    if len(args) == 0:
        return []
    var_in = args[0] if len(args) == 1 else args
    if var_in is None:
        return []
    if isinstance(var_in, list):
        return var_in
    if isinstance(var_in, (tuple, set)):
        return list(var_in)
    return [var_in]


def _caller_location() -> str:
    for fr in inspect.stack():
        if fr.filename != __file__:
            return f'File "{fr.filename}", line {fr.lineno}, in {fr.function}'
    return "N/A"


class Signal:
    """State/sensitivity carrier (``core_objects.py:71-195``)."""

    def __init__(self, tag: str = "", state: Any = None, sensitivity: Any = None, min: Any = None, max: Any = None):
        self.tag = tag
        self.state = state
        self.sensitivity = sensitivity
        self.min = min
        self.max = max
        self.keep_alloc = sensitivity is not None
        self._init_loc = _caller_location()

    def _err_str(self):
        return f"\n\t| Signal '{self.tag}', initialized in {self._init_loc}"

    def add_sensitivity(self, ds: Any):
        """Accumulate a sensitivity term: deep copy on first use, in-place add afterwards."""
        try:
            if ds is None:
                return
            if self.sensitivity is None:
                self.sensitivity = copy.deepcopy(ds)
            elif hasattr(self.sensitivity, "add_sensitivity"):
                self.sensitivity.add_sensitivity(ds)
            else:
                self.sensitivity += ds
            return self
        except TypeError:
            if isinstance(ds, type(self.sensitivity)):
                raise TypeError(f"Cannot add to the sensitivity with type '{type(self.sensitivity).__name__}'"
                                + self._err_str())
            raise TypeError(f"Adding wrong type '{type(ds).__name__}' to the sensitivity "
                            f"'{type(self.sensitivity).__name__}'" + self._err_str())
        except ValueError:
            s1 = getattr(self.sensitivity, "shape", ())
            s2 = getattr(ds, "shape", ())
            raise ValueError(f"Cannot add argument of shape {s2} to the sensitivity of shape {s1}"
                             + self._err_str()) from None

    def reset(self, keep_alloc: bool = None):
        """Zero the sensitivity (keeping its allocation) or drop it."""
        if self.sensitivity is None:
            return self
        if keep_alloc is None:
            keep_alloc = self.keep_alloc
        if keep_alloc:
            try:
                self.sensitivity[...] = 0
            except TypeError:
                try:
                    self.sensitivity *= 0
                except TypeError:
                    self.sensitivity = None
        else:
            self.sensitivity = None
        return self

    def __getitem__(self, item):
        return SignalSlice(self, item)

    def __repr__(self):
        st = "empty state" if self.state is None else f"state of type {type(self.state).__name__}"
        return f'Signal "{self.tag}" with {st}'


class SignalSlice(Signal):
    """A view into part of a base signal (``core_objects.py:198-315``)."""

    def __init__(self, base, sl, tag=None):
        self.base = base
        self.slice = sl
        self.keep_alloc = False
        self.tag = f"{base.tag}[{sl}]" if tag is None else tag
        self._init_loc = _caller_location()

    @property
    def state(self):
        return None if self.base.state is None else self.base.state[self.slice]

    @state.setter
    def state(self, new_state):
        self.base.state[self.slice] = new_state

    @property
    def sensitivity(self):
        return None if self.base.sensitivity is None else self.base.sensitivity[self.slice]

    @sensitivity.setter
    def sensitivity(self, new_sens):
        if self.base.sensitivity is None:
            if new_sens is None:
                return
            self.base.sensitivity = self.base.state * 0
        self.base.sensitivity[self.slice] = 0 if new_sens is None else new_sens

    def add_sensitivity(self, ds: Any):
        if ds is None:
            return
        if self.base.sensitivity is None:
            self.base.sensitivity = self.base.state * 0
        self.base.sensitivity[self.slice] += ds
        return self

    def reset(self, keep_alloc: bool = None):
        if self.sensitivity is not None:
            self.sensitivity = None
        return self


def make_signals(*args: str, **kwargs: Any):
    """Batch-create signals by tag (``core_objects.py:318-357``)."""
    ret_args = tuple(Signal(tag) for tag in args)
    ret_kwargs = {tag: Signal(tag, val) for tag, val in kwargs.items()}
    if args and kwargs:
        return ret_args, ret_kwargs
    if args:
        return ret_args[0] if len(ret_args) == 1 else ret_args
    return ret_kwargs


def _is_valid_signal(sig: Any) -> bool:
    return isinstance(sig, Signal) or all(hasattr(sig, f) for f in ("state", "sensitivity", "add_sensitivity", "reset"))


def _is_valid_module(mod: Any) -> bool:
    return isinstance(mod, Module) or all(hasattr(mod, f) for f in ("response", "sensitivity", "reset"))


def _check_function_signature(fn, n_args: int = None):
    """Positional-only contract for ``__call__``/``_sensitivity`` (``core_objects.py:408-448``)."""
    min_args, max_args = 0, 0
    name = f"{type(getattr(fn, '__self__', None)).__name__}.{fn.__name__}"
    for p in inspect.signature(fn).parameters.values():
        if p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD):
            if p.default is p.empty:
                min_args += 1
            if max_args is not None:
                max_args += 1
        elif p.kind == p.VAR_POSITIONAL:
            max_args = None
        elif p.kind == p.KEYWORD_ONLY:
            raise SyntaxError(f'{name} may not contain keyword-only arguments "{p}"')
        elif p.kind == p.VAR_KEYWORD:
            raise SyntaxError(f'{name} may not contain dict of keyword-only arguments "{p}"')
    if n_args is not None:
        if n_args < min_args:
            raise TypeError(f"Not enough arguments ({n_args}) for {name}; expected {min_args}")
        if max_args is not None and n_args > max_args:
            raise TypeError(f"Too many arguments ({n_args}) for {name}; expected {max_args}")
    return min_args, max_args


class Module(ABC):
    """Base class: ``__call__`` computes the response, ``_sensitivity`` its adjoint."""

    sig_in: List = None
    sig_out: List = None
    _init_loc: str = None

    def __init_subclass__(cls, **kwargs):
        super().__init_subclass__(**kwargs)
        fwd = cls.__dict__.get("__call__", None)
        if fwd is None or getattr(fwd, "__isabstractmethod__", False) or getattr(fwd, "_is_wrapped", False):
            return
        _check_function_signature(fwd)
        cls._orig_call = fwd
        cls._response = cls._wrap_connect(fwd, create=False)
        cls.__call__ = cls._wrap_connect(fwd, create=True)

    @classmethod
    def _wrap_connect(cls, response, create=True):
        def wrapped(self, *args):
            inp = [s.state if _is_valid_signal(s) else s for s in args]
            all_constant = len(args) > 0 and not any(_is_valid_signal(s) for s in args)
            self._init_loc = _caller_location()
            if create and not all_constant:
                self.sig_in = _parse_to_list(args)
            out = response(self, *inp)
            if all_constant:
                return out
            state_out = _parse_to_list(out)
            if create:
                if self.sig_out is None:
                    self.sig_out = [Signal(f"{type(self).__name__}_output{i}") for i in range(len(state_out))]
                _check_function_signature(self._sensitivity, self.n_out)
                for n in Network.active:
                    if self in n.mods:
                        raise RuntimeError("Module already in network, cannot add twice.")
                    n.append(self)
            if len(state_out) != self.n_out:
                raise TypeError(f"Number of responses calculated ({len(state_out)}) is unequal to number of "
                                f"output signals ({self.n_out})")
            for i, val in enumerate(state_out):
                self.sig_out[i].state = val
            if self.n_out == 0:
                return None
            return self.sig_out[0] if self.n_out == 1 else tuple(self.sig_out)
        wrapped._is_wrapped = True
        return wrapped

    def connect(self, sig_in, sig_out=None):
        """Connect signals without adding to an active network."""
        self.sig_in = _parse_to_list(sig_in)
        inp = [s.state if _is_valid_signal(s) else s for s in self.sig_in]
        out = _parse_to_list(self._orig_call(*inp))
        if sig_out is not None:
            self.sig_out = _parse_to_list(sig_out)
            if self.n_out != len(out):
                raise TypeError(f"Number of responses calculated ({len(out)}) is unequal to number of "
                                f"output signals ({self.n_out})")
        else:
            self.sig_out = [Signal(f"{type(self).__name__}_output{i}") for i in range(len(out))]
        _check_function_signature(self._sensitivity, self.n_out)
        for i, val in enumerate(out):
            self.sig_out[i].state = val
        return self

    @property
    def n_in(self):
        return None if self.sig_in is None else len(self.sig_in)

    @property
    def n_out(self):
        return None if self.sig_out is None else len(self.sig_out)

    def _err_str(self):
        ins = "Unconnected" if self.sig_in is None else ", ".join(str(getattr(s, "tag", type(s).__name__))
                                                                  for s in self.sig_in)
        outs = "Unconnected" if self.sig_out is None else ", ".join(str(getattr(s, "tag", type(s).__name__))
                                                                    for s in self.sig_out)
        return f"\n\t| Module '{type(self).__name__}'( {ins} ) --> {outs}\n\t| Used in {self._init_loc}"

    def response(self):
        """Recompute outputs from the current input states."""
        try:
            self._response(*self.sig_in)
            return self
        except Exception as e:
            raise type(e)(str(e) + "\n\t| Above error was raised when calling response(). Module details:"
                          + self._err_str()).with_traceback(sys.exc_info()[2])

    def sensitivity(self):
        """Back-propagate output sensitivities into the input signals."""
        try:
            sens_in = self.get_output_sensitivities(as_list=True)
            has_in = self.sig_in is not None and any(_is_valid_signal(s) for s in self.sig_in)
            has_out = self.sig_out is not None and any(_is_valid_signal(s) for s in self.sig_out)
            if not (has_in or has_out):
                return
            if self.n_out > 0 and all(s is None for s in sens_in):
                return
            sens_out = _parse_to_list(self._sensitivity(*sens_in))
            if len(sens_out) != self.n_in:
                raise TypeError(f"Number of sensitivities calculated ({len(sens_out)}) is unequal to number of "
                                f"input signals ({self.n_in})")
            for i, ds in enumerate(sens_out):
                if _is_valid_signal(self.sig_in[i]):
                    self.sig_in[i].add_sensitivity(ds)
            return self
        except Exception as e:
            raise type(e)(str(e) + "\n\t| Above error was raised when calling sensitivity(). Module details:"
                          + self._err_str()).with_traceback(sys.exc_info()[2])

    def reset(self):
        """Reset sensitivities of all connected signals."""
        for s in self.sig_out or []:
            if _is_valid_signal(s):
                s.reset()
        for s in self.sig_in or []:
            if _is_valid_signal(s):
                s.reset()
        self._reset()
        return self

    def get_input_states(self, as_list=False):
        if self.sig_in is None:
            return None
        if not as_list and self.n_in == 1:
            return self.sig_in[0].state
        return [s.state if _is_valid_signal(s) else s for s in self.sig_in]

    def get_output_states(self, as_list=False):
        if self.sig_out is None:
            return None
        if not as_list and self.n_out == 1:
            return self.sig_out[0].state
        return [s.state if _is_valid_signal(s) else s for s in self.sig_out]

    def get_output_sensitivities(self, as_list=False):
        if self.sig_out is None:
            return None
        if not as_list and self.n_out == 1:
            return self.sig_out[0].sensitivity
        return [s.sensitivity if _is_valid_signal(s) else s for s in self.sig_out]

    @abstractmethod
    def __call__(self, *args):
        raise NotImplementedError("No response behavior defined")

    def _sensitivity(self, *args):
        if self.n_out and self.n_in:
            warnings.warn(f"Sensitivity routine is used, but not defined, in {type(self).__name__}", RuntimeWarning)
        return [None for _ in self.sig_in]

    def _reset(self):
        pass

    def __repr__(self):
        return f"Module {type(self).__name__} at {hex(id(self))}"


class Network(Module):
    """Ordered list of modules run as one (``core_objects.py:750-994``)."""

    active = []

    def __init__(self, *args, print_timing=False):
        self._init_loc = _caller_location()
        self.mods = []
        self.append(*args)
        self.print_timing = print_timing

    def __enter__(self):
        if self in Network.active:
            raise ValueError("Network is already activated, cannot activate twice")
        Network.active.append(self)
        return self

    def __exit__(self, typ, value, traceback):
        Network.active.remove(self)

    def _run(self, fns, label):
        if self.print_timing is False:
            for fn, _ in fns:
                fn()
            return
        tmin = 0.0 if isinstance(self.print_timing, bool) else float(self.print_timing)
        t0 = time.time()
        for fn, name in fns:
            t1 = time.time()
            fn()
            if time.time() - t1 > tmin:
                print(f'-- {label} of "{name}" took {time.time() - t1} s')
        if time.time() - t0 > tmin:
            print(f"-- TOTAL {label} took {time.time() - t0} s")

    def response(self):
        self._run([(m.response, type(m).__name__) for m in self.mods], "Response")

    def sensitivity(self):
        self._run([(m.sensitivity, type(m).__name__) for m in reversed(self.mods)], "Sensitivity")

    def reset(self):
        for m in reversed(self.mods):
            m.reset()

    def __len__(self):
        return len(self.mods)

    def __iter__(self):
        return iter(self.mods)

    def __getitem__(self, sel):
        out = Network(self.mods[sel])
        if len(out) == 0:
            return None
        return out.mods[0] if len(out) == 1 else out

    def __call__(self, *args):
        raise RuntimeError("Cannot re-connect signals in existing network")

    def append(self, *newmods):
        modlist = _parse_to_list(*newmods)
        if not modlist:
            return None
        for i, m in enumerate(modlist):
            if not _is_valid_module(m):
                raise TypeError(f"Argument #{i} is not a valid module, type='{type(m).__name__}'.")
        self.mods.extend(modlist)
        so = modlist[-1].sig_out or []
        return so[0] if len(so) == 1 else so

    @staticmethod
    def _signal_set(sigs):
        def base(s):
            return base(s.base) if isinstance(s, SignalSlice) else s
        return set(base(s) for s in _parse_to_list(sigs) if _is_valid_signal(s))

    @property
    def sig_in(self):
        ins, outs = set(), set()
        for m in self.mods:
            ins |= self._signal_set(m.sig_in)
            outs |= self._signal_set(m.sig_out)
        return list(ins - outs)

    @property
    def sig_out(self):
        ins, outs = set(), set()
        for m in self.mods:
            ins |= self._signal_set(m.sig_in)
            outs |= self._signal_set(m.sig_out)
        return list(outs - ins)

    def get_input_cone(self, fromsig=None, frommod=None):
        touched = self._signal_set(fromsig)
        frommod = set(_parse_to_list(frommod))
        if not touched and not frommod:
            return self
        cone = Network()
        for m in self:
            if m in frommod or any(s in touched for s in self._signal_set(m.sig_in)):
                touched |= self._signal_set(m.sig_out)
                cone.append(m)
        return cone

    def get_output_cone(self, tosig=None, tomod=None):
        dep = self._signal_set(tosig)
        tomod = set(_parse_to_list(tomod))
        if not dep and not tomod:
            return self
        cone = []
        for m in reversed(self.mods):
            if m in tomod or any(s in dep for s in self._signal_set(m.sig_out)):
                dep |= self._signal_set(m.sig_in)
                cone.append(m)
        return Network(list(reversed(cone)))


Network.active.append(Network())
