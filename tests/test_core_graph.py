"""The Signal / Module / Network contract on the host (no GPU).

The cases follow what pyMOTO's own tests require of its core objects
(``tests/test_core_signal.py``, ``tests/test_core_module.py``, ``tests/test_core_network.py`` of the
reference): sensitivity accumulation and reset, signal slices, auto-connection of modules, the
positional-only rule, skipped sensitivities, reconnection errors, and the network cones.
"""
import copy

import numpy as np
import numpy.testing as npt
import pytest

import pymoto_amd as pm


class Scale(pm.Module):
    def __init__(self, factor=2.0):
        self.factor = factor

    def __call__(self, x):
        return self.factor * x

    def _sensitivity(self, dy):
        return self.factor * dy


class MulAdd(pm.Module):
    def __init__(self):
        self.evaluations = 0
        self.adjoint_ran = False

    def __call__(self, a, b):
        self.evaluations += 1
        return a * b, a + b

    def _sensitivity(self, dprod, dsum):
        self.adjoint_ran = True
        a, b = self.get_input_states()
        dprod = 0.0 if dprod is None else dprod
        dsum = 0.0 if dsum is None else dsum
        return b * dprod + dsum, a * dprod + dsum

    def _reset(self):
        self.evaluations = 0
        self.adjoint_ran = False


class Add(pm.Module):
    def __call__(self, a, b):
        return a + b

    def _sensitivity(self, dc):
        return dc, dc


class Source(pm.Module):
    def __init__(self, value=1.5):
        self.value = value

    def __call__(self):
        return self.value


class Sink(pm.Module):
    def __call__(self, x):
        self.seen = x

    def _sensitivity(self):
        return 0.5


class TwoOut(pm.Module):
    def __call__(self, x):
        return 2 * x, 3 * x


class Double(pm.Module):
    def __call__(self, x):
        return 2 * x


# ---------------------------------------------------------------------------------------------
# signals

def test_signal_accumulates_and_resets():
    s = pm.Signal("s")
    assert s.state is None and s.sensitivity is None and s.tag == "s"
    s.sensitivity = 2.0
    s.add_sensitivity(3.0)
    assert s.sensitivity == 5.0
    s.reset(keep_alloc=True)
    assert s.sensitivity == 0.0
    s.reset()
    assert s.sensitivity is None
    v = np.array([1.0, 2.0, 3.0])
    s.add_sensitivity(v)
    assert not np.shares_memory(s.sensitivity, v)  # first term is copied
    s.add_sensitivity(None)
    npt.assert_equal(s.sensitivity, v)
    s.add_sensitivity(v)
    npt.assert_equal(s.sensitivity, 2 * v)
    k = pm.Signal("k", np.zeros(3), np.ones(3))
    assert k.keep_alloc
    mem = k.sensitivity
    k.reset()
    assert k.sensitivity is mem and not np.any(mem)


def test_signal_add_errors_keep_their_type():
    s = pm.Signal("s", sensitivity=np.ones(3))
    with pytest.raises(TypeError):
        s.add_sensitivity("text")
    with pytest.raises(ValueError, match="signal 's'"):
        s.add_sensitivity(np.ones((3, 3)))
    r = pm.Signal("r", np.ones(3))
    r.add_sensitivity(np.ones(3) + 1j * np.ones(3))  # complex onto an empty sensitivity is a copy

    class Opaque:
        pass
    o = pm.Signal("o")
    o.add_sensitivity(Opaque())
    with pytest.raises(TypeError):
        o.add_sensitivity(Opaque())
    with pytest.warns(RuntimeWarning):
        o.reset(True)
    assert o.sensitivity is None


def test_make_signals_forms():
    a = pm.make_signals("a")
    assert isinstance(a, pm.Signal) and a.tag == "a"
    a, b = pm.make_signals("a", "b")
    assert (a.tag, b.tag) == ("a", "b")
    d = pm.make_signals(x=1.0, y=2.0)
    assert d["x"].state == 1.0 and d["y"].tag == "y"
    plain, named = pm.make_signals("p", q=3)
    assert isinstance(plain, tuple) and plain[0].tag == "p" and named["q"].state == 3


def test_signal_slices_read_and_write_the_base():
    s = pm.Signal("x", np.arange(10, dtype=float))
    assert s[3].state == 3.0
    s[2:5].state = 7.0
    npt.assert_equal(s.state[2:5], 7.0)
    s[0].state += 1.0
    assert s.state[0] == 1.0
    assert s[1].sensitivity is None
    s[1].add_sensitivity(1.0)
    assert s.sensitivity.shape == s.state.shape and s[0].sensitivity == 0.0
    with pytest.raises(ValueError):
        s[1].add_sensitivity(np.array([1.0, 2.0]))
    s[4:7].add_sensitivity(np.array([1.0, 2.0, 3.0]))
    s[4:6].reset()
    npt.assert_equal(s.sensitivity[4:7], [0.0, 0.0, 3.0])
    assert s[1].sensitivity == 1.0
    s[np.arange(7, 9)].add_sensitivity(np.array([5.0, 6.0]))
    npt.assert_equal(s.sensitivity[7:9], [5.0, 6.0])
    m = pm.Signal("m", np.zeros((4, 4)))
    m[0, 1:3].sensitivity = 1.0
    assert m.sensitivity.shape == (4, 4) and m.sensitivity.sum() == 2.0
    with pytest.raises(IndexError):
        _ = m[1, 2, 3].state
    with pytest.raises(TypeError):
        pm.Signal("empty")[2].sensitivity = 1.0


# ---------------------------------------------------------------------------------------------
# modules

def test_module_is_abstract_without_a_response():
    with pytest.raises(TypeError):
        pm.Module()

    class NoResponse(pm.Module):
        def _sensitivity(self, d):
            return d
    with pytest.raises(TypeError):
        NoResponse()


def test_module_connects_and_evaluates():
    a = pm.Signal("a", 1.5)
    m = Scale()
    b = m(a)
    assert isinstance(b, pm.Signal) and b.tag == "Scale_output0" and b.state == 3.0
    assert m.sig_in == [a] and m.sig_out == [b]
    assert m(4.0) == 8.0  # plain values: a function call
    c = Scale()
    assert c(4.0) == 8.0 and c.sig_in is None and c.sig_out is None


def test_mixed_constant_and_signal_inputs():
    a = pm.Signal("a", 2.0)
    m = MulAdd()
    p, s = m(a, 3.0)
    assert m.sig_in == [a, 3.0] and (p.state, s.state) == (6.0, 5.0)
    p.sensitivity = 1.0
    m.sensitivity()
    assert a.sensitivity == 3.0


def test_response_sensitivity_reset_cycle():
    sa, sb = pm.Signal("a", 2.5), pm.Signal("b", 3.5)
    m = MulAdd()
    sc, sd = m(sa, sb)
    assert m.evaluations == 1
    m.response()
    assert m.evaluations == 2
    m.sensitivity()
    assert not m.adjoint_ran  # every output sensitivity is None
    sc.sensitivity = 1.0
    m.sensitivity()
    assert m.adjoint_ran and sa.sensitivity == 3.5 and sb.sensitivity == 2.5
    m.reset()
    assert all(s.sensitivity is None for s in (sa, sb, sc, sd)) and m.evaluations == 0
    sc.sensitivity = 1.0
    sd.sensitivity = 1.0
    m.sensitivity()
    assert (sa.sensitivity, sb.sensitivity) == (4.5, 3.5)


def test_source_and_sink_modules():
    src = Source(3.25)
    s = src()
    assert isinstance(s, pm.Signal) and s.state == 3.25
    src.value = 4.0
    src.response()
    assert s.state == 4.0
    a = pm.Signal("a", 1.0)
    sink = Sink()
    assert sink(a) is None and sink.seen == 1.0
    sink.sensitivity()
    sink.sensitivity()
    assert a.sensitivity == 1.0  # accumulated twice
    sink.reset()
    assert a.sensitivity is None
    k = Sink()
    k(2.0)  # constant input: nothing is wired, sensitivity is a no-op
    k.sensitivity()


def test_counts_are_checked():
    a = pm.Signal("a", 1.0)
    with pytest.raises(TypeError):
        Scale()(a, a)  # one parameter, two arguments

    class OneOfTwo(pm.Module):
        def __call__(self, x, y):
            return x * y

        def _sensitivity(self, dz):
            return dz
    n = OneOfTwo().connect([pm.Signal("x", 2.0), pm.Signal("y", 3.0)])
    n.sig_out[0].sensitivity = 1.0
    with pytest.raises(TypeError, match="OneOfTwo"):
        n.sensitivity()
    assert n.n_out == 1 and n.n_in == 2


def test_positional_only_rule():
    with pytest.raises(SyntaxError):
        class KwOnly(pm.Module):
            def __call__(self, x, *, scale):
                return x
    with pytest.raises(SyntaxError):
        class VarKw(pm.Module):
            def __call__(self, x, **kw):
                return x

    class Defaults(pm.Module):
        def __call__(self, a, b=1.0):
            return a + b

        def _sensitivity(self, dc):
            return dc if self.n_in == 1 else (dc, dc)
    sa, sb = pm.Signal("a", np.array([1.0, 2.0])), pm.Signal("b", np.array([3.0, 4.0]))
    m = Defaults()
    c = m(sa)
    npt.assert_equal(c.state, [2.0, 3.0])
    c.sensitivity = np.ones(2)
    m.sensitivity()
    npt.assert_equal(sa.sensitivity, [1.0, 1.0])
    with pytest.raises(TypeError):
        Defaults()(sa, b=sb)


def test_identical_sensitivities_are_not_shared():
    sa, sb = pm.Signal("a", np.ones(3)), pm.Signal("b", np.ones(3))
    m = Add()
    sc = m(sa, sb)
    sc.sensitivity = np.ones(3)
    m.sensitivity()
    assert not np.may_share_memory(sa.sensitivity, sb.sensitivity)
    assert not np.may_share_memory(sa.sensitivity, sc.sensitivity)


def test_errors_carry_module_context():
    class Fails(pm.Module):
        def __call__(self, x):
            return x

        def _sensitivity(self, dy):
            raise ValueError("bad adjoint")
    s = pm.Signal("s", 1.0)
    m = Fails()
    y = m(s)
    y.sensitivity = 1.0
    with pytest.raises(ValueError, match="bad adjoint(.|\n)*Fails"):
        m.sensitivity()


def test_a_module_is_connected_once():
    sa, sb = pm.Signal("a", np.ones(2)), pm.Signal("b", np.ones(2))
    m = Add()
    se = m(sa, sb)
    with pytest.raises(RuntimeError):
        m(sb, sa)
    npt.assert_equal(se.state, [2.0, 2.0])
    with pm.Network() as fn:
        m2 = Add()
        m2(sa, sb)
        sf = copy.deepcopy(m2)(sb, sb)  # a copy is a new module
    assert len(fn) == 2 and fn[1] is not m2
    npt.assert_equal(sf.state, [2.0, 2.0])
    mc = Add().connect([sa, sb])
    assert isinstance(mc, Add) and mc.sig_out[0].state[0] == 2.0


# ---------------------------------------------------------------------------------------------
# networks

def test_network_chain_rule():
    x1, x2 = pm.Signal("x1", 2.0), pm.Signal("x2", 3.0)
    with pm.Network() as fn:
        y1 = Scale(2.0)(x1)
        y2 = Scale(5.0)(x2)
        z, _ = MulAdd()(y1, y2)
    fn.response()
    assert (y1.state, y2.state, z.state) == (4.0, 15.0, 60.0)
    z.sensitivity = 1.0
    fn.sensitivity()
    assert (y1.sensitivity, y2.sensitivity) == (15.0, 4.0)
    assert (x1.sensitivity, x2.sensitivity) == (30.0, 20.0)
    fn.reset()
    assert all(s.sensitivity is None for s in (x1, x2, y1, y2, z))
    with pytest.raises(RuntimeError):
        fn(x1)
    with pytest.raises(ValueError):
        with fn:
            with fn:
                pass


def test_network_append_and_duck_typing():
    class Foreign:  # a module of another library: only the three methods
        sig_in, sig_out = [], []

        def response(self):
            pass

        def sensitivity(self):
            pass

        def reset(self):
            pass
    fn = pm.Network()
    fn.append(Foreign())
    assert len(fn) == 1
    with pytest.raises(TypeError):
        fn.append(object())


def _graph():
    """The reference's subset test graph: two chains and a detached source/sink pair."""
    sx, sy = pm.Signal("x", 1.0), pm.Signal("y", 2.0)
    with pm.Network() as fn:
        m = {}
        for k in ("m1",):
            m[k] = TwoOut()
        for k in ("m2", "m2a", "m5", "m9"):
            m[k] = Double()
        m["m3"], m["m4"] = Sink(), Source()
        for k in ("m6", "m7", "m8", "m10"):
            m[k] = Add()
        sx1, sx2 = m["m1"](sx)
        sy1 = m["m2"](sy)
        sy1a = m["m2a"](sy1)
        m["m3"](sy1a)
        sz = m["m4"]()
        sz1 = m["m5"](sz)
        so1 = m["m6"](sz1, sx2)
        so2 = m["m7"](sy1, sx1)
        sg1 = m["m8"](so1, so2)
        sg2 = m["m9"](sy1)
        sg3 = m["m10"](sg1, 2.5)
        m["d1"], m["d2"], m["d3"] = Source(), Double(), Sink()
        sd1 = m["d1"]()
        sd2 = m["d2"](sd1)
        m["d3"](sd2)
    sig = dict(x=sx, y=sy, x1=sx1, x2=sx2, y1=sy1, o1=so1, o2=so2, g1=sg1, g2=sg2, g3=sg3, d1=sd1, d2=sd2)
    return fn, m, sig


def _names(net, m):
    inv = {id(v): k for k, v in m.items()}
    return sorted(inv[id(mod)] for mod in net)


def test_network_cones_and_slices():
    fn, m, s = _graph()
    every = sorted(m)
    cases = [
        (fn.get_input_cone(s["d1"]), ["d2", "d3"]),
        (fn.get_input_cone(s["o1"]), ["m10", "m8"]),
        (fn.get_input_cone(frommod=m["d1"]), ["d1", "d2", "d3"]),
        (fn.get_input_cone([s["y1"], s["o1"]]), ["m10", "m2a", "m3", "m7", "m8", "m9"]),
        (fn.get_input_cone(), every),
        (fn.get_output_cone(s["d2"]), ["d1", "d2"]),
        (fn.get_output_cone(s["o1"]), ["m1", "m4", "m5", "m6"]),
        (fn.get_output_cone(tomod=m["d3"]), ["d1", "d2", "d3"]),
        (fn.get_output_cone([s["o2"], s["o1"]]), ["m1", "m2", "m4", "m5", "m6", "m7"]),
        (fn.get_output_cone(tosig=set()), every),
    ]
    for net, want in cases:
        assert _names(net, m) == sorted(want)
    assert fn[0] is m["m1"] and fn[-1] is m["d3"]
    assert fn[s["d1"]:s["d2"]] is m["d2"]
    assert _names(fn[1:4], m) == ["m2", "m2a", "m3"]
    assert _names(fn[s["d1"]:], m) == ["d2", "d3"]
    assert _names(fn[:s["d2"]], m) == ["d1", "d2"]


@pytest.mark.parametrize("sinks", [True, False])
@pytest.mark.parametrize("sources", [True, False])
def test_network_subsets(sinks, sources):
    fn, m, s = _graph()
    snk = ["m2a", "m3"] if sinks else []
    src = ["m4", "m5"] if sources else []
    kw = dict(include_sinks=sinks, include_sources=sources)
    cases = [
        (fn.get_subset(fromsig=s["x"], **kw), ["m1", "m6", "m7", "m8", "m10"] + src),
        (fn.get_subset(fromsig=[s["x"], s["y"]], **kw), ["m1", "m2", "m6", "m7", "m8", "m9", "m10"] + src + snk),
        (fn.get_subset(fromsig=s["x2"], **kw), ["m6", "m8", "m10"] + src),
        (fn.get_subset(tosig=s["g1"], **kw), ["m1", "m2", "m6", "m7", "m8"] + src + snk),
        (fn.get_subset(tosig=[s["g2"], s["x2"]], **kw), ["m1", "m2", "m9"] + snk),
        (fn.get_subset(tosig=s["g3"], **kw), ["m1", "m2", "m6", "m7", "m8", "m10"] + src + snk),
        (fn.get_subset(fromsig=s["d1"], tosig=s["d2"], **kw),
         ["d2"] + (["d1"] if sources else []) + (["d3"] if sinks else [])),
    ]
    for net, want in cases:
        assert _names(net, m) == sorted(want)


def test_network_cones_with_slices_and_constants():
    sx = pm.Signal("x", np.array([1.0, 2.0, 3.0]))
    with pm.Network() as fn:
        d1, d2, d3, d4 = Source(), Add(), Add(), Sink()
        s1 = d1()
        s2 = d2(s1, np.array([2.0]))  # an unhashable constant input
        s3 = d3(s2, sx[2])
        d4(s3)
    m = dict(d1=d1, d2=d2, d3=d3, d4=d4)
    assert _names(fn.get_input_cone(sx), m) == ["d3", "d4"]
    assert _names(fn.get_output_cone(s2), m) == ["d1", "d2"]
    assert _names(fn.get_subset(sx[1], s3[0], include_sinks=False, include_sources=False), m) == ["d3"]
    assert _names(fn.get_subset(sx, s3, include_sinks=True, include_sources=True), m) == ["d1", "d2", "d3", "d4"]


def test_network_timing_prints(capsys):
    x = pm.Signal("x", 1.0)
    with pm.Network(print_timing=True) as fn:
        Scale()(x)
    fn.response()
    assert "response of Scale" in capsys.readouterr().out
