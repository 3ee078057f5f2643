"""Oracle of the learned warm start's network (test infrastructure only).

A numpy restatement of the encoder-decoder in ``pymoto_amd/learned.py`` (no reference counterpart:
the learned warm start is new).  Arrays are channels-last (nz, ny, nx, c); weights W[tap][cout][cin]
with tap = (dz+1)*9 + (dy+1)*3 + (dx+1); 3x3x3 convolutions with zero padding; 2x2x2 average
pooling; nearest upsampling plus skip; nodal output = mean of channels 0..2 over the elements
around each node.  ``bf16=True`` rounds every activation and activation gradient to bfloat16
(round to nearest even) where the device does, so the comparison isolates accumulation order.

Pinning: the backward pass (every weight gradient) is checked against central finite differences
of the FP64 forward pass (tests/test_learned_oracle.py); the convolution against a direct
triple-loop evaluation on a tiny grid.
"""
import numpy as np

SPEC = [
    ("e0a", 0, 8, 16, True), ("e0b", 0, 16, 16, True),
    ("e1a", 1, 16, 32, True), ("e1b", 1, 32, 32, True),
    ("e2a", 2, 32, 64, True), ("e2b", 2, 64, 64, True),
    ("b0", 3, 64, 64, True), ("b1", 3, 64, 64, True),
    ("d2a", 2, 64, 64, True), ("d2b", 2, 64, 32, True),
    ("d1a", 1, 32, 32, True), ("d1b", 1, 32, 16, True),
    ("d0a", 0, 16, 16, True), ("d0b", 0, 16, 16, False),
]
TAPS = [(t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1) for t in range(27)]  # (dx, dy, dz)


def bf16_round(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(X, dx, dy, dz):
    """S[z, y, x] = X[z+dz, y+dy, x+dx] (zero outside)."""
    nz, ny, nx = X.shape[:3]
    P = np.zeros((nz + 2, ny + 2, nx + 2) + X.shape[3:])
    P[1:-1, 1:-1, 1:-1] = X
    return P[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]


def conv_fwd(X, W, b=None, relu=False):
    Y = np.zeros(X.shape[:3] + (W.shape[1],))
    for t, (dx, dy, dz) in enumerate(TAPS):
        Y += _shift(X, dx, dy, dz) @ W[t].T
    if b is not None:
        Y += b
    return np.maximum(Y, 0.0) if relu else Y


def conv_bwd_data(G, W):
    dX = np.zeros(G.shape[:3] + (W.shape[2],))
    for t, (dx, dy, dz) in enumerate(TAPS):
        dX += _shift(G, -dx, -dy, -dz) @ W[t]
    return dX


def conv_bwd_weight(X, G):
    c_out, c_in = G.shape[3], X.shape[3]
    dW = np.zeros((27, c_out, c_in))
    g2 = G.reshape(-1, c_out)
    for t, (dx, dy, dz) in enumerate(TAPS):
        dW[t] = g2.T @ _shift(X, dx, dy, dz).reshape(-1, c_in)
    return dW, g2.sum(0)


def pool(X):
    nz, ny, nx, c = X.shape
    return X.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool_bwd(G):
    return np.repeat(np.repeat(np.repeat(G, 2, 0), 2, 1), 2, 2) / 8.0


def up(X):
    return np.repeat(np.repeat(np.repeat(X, 2, 0), 2, 1), 2, 2)


def up_bwd(G):
    nz, ny, nx, c = G.shape
    return G.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


def node_counts(ez, ey, ex):
    cz = np.full(ez + 1, 2.0)
    cy = np.full(ey + 1, 2.0)
    cx = np.full(ex + 1, 2.0)
    cz[[0, -1]] = cy[[0, -1]] = cx[[0, -1]] = 1.0
    return cz[:, None, None] * cy[None, :, None] * cx[None, None, :]


def elem_to_node(Xe):
    ez, ey, ex, _ = Xe.shape
    U = np.zeros((ez + 1, ey + 1, ex + 1, 3))
    for a in range(8):
        ox, oy, oz = a & 1, (a >> 1) & 1, (a >> 2) & 1
        U[oz:oz + ez, oy:oy + ey, ox:ox + ex] += Xe[..., :3]
    return (U / node_counts(ez, ey, ex)[..., None]).reshape(-1)


def elem_to_node_bwd(dU, shape, c):
    ez, ey, ex = shape
    D = dU.reshape(ez + 1, ey + 1, ex + 1, 3) / node_counts(ez, ey, ex)[..., None]
    G = np.zeros((ez, ey, ex, c))
    for a in range(8):
        ox, oy, oz = a & 1, (a >> 1) & 1, (a >> 2) & 1
        G[..., :3] += D[oz:oz + ez, oy:oy + ey, ox:ox + ex]
    return G


class Net:
    """params: name -> (W (27, cout, cin), b (cout,)); activations kept for the backward pass."""

    def __init__(self, params, bf16=False):
        self.p = params
        self.rnd = bf16_round if bf16 else (lambda a: a)

    def _conv(self, name, X):
        _, _, _, _, relu = next(s for s in SPEC if s[0] == name)
        W, b = self.p[name]
        return self.rnd(conv_fwd(X, W, b, relu))

    def forward(self, X):
        r, a = self.rnd, {}
        self.a = a
        a["in"] = X
        a["e0a"] = self._conv("e0a", X)
        a["e0b"] = self._conv("e0b", a["e0a"])
        a["p0"] = r(pool(a["e0b"]))
        a["e1a"] = self._conv("e1a", a["p0"])
        a["e1b"] = self._conv("e1b", a["e1a"])
        a["p1"] = r(pool(a["e1b"]))
        a["e2a"] = self._conv("e2a", a["p1"])
        a["e2b"] = self._conv("e2b", a["e2a"])
        a["p2"] = r(pool(a["e2b"]))
        a["b0"] = self._conv("b0", a["p2"])
        a["b1"] = self._conv("b1", a["b0"])
        a["u2"] = r(up(a["b1"]) + a["e2b"])
        a["d2a"] = self._conv("d2a", a["u2"])
        a["d2b"] = self._conv("d2b", a["d2a"])
        a["u1"] = r(up(a["d2b"]) + a["e1b"])
        a["d1a"] = self._conv("d1a", a["u1"])
        a["d1b"] = self._conv("d1b", a["d1a"])
        a["u0"] = r(up(a["d1b"]) + a["e0b"])
        a["d0a"] = self._conv("d0a", a["u0"])
        a["out"] = self._conv("d0b", a["d0a"])
        return elem_to_node(a["out"])

    def backward(self, dU):
        a, r = self.a, self.rnd
        grads = {}

        def layer(name, x_in, g, relu_out=None, need_dx=True):
            if relu_out is not None:
                g = g * (relu_out > 0)
            grads[name] = conv_bwd_weight(x_in, g)
            return r(conv_bwd_data(g, self.p[name][0])) if need_dx else None

        g = r(elem_to_node_bwd(dU, a["out"].shape[:3], a["out"].shape[3]))
        g = layer("d0b", a["d0a"], g)
        g = layer("d0a", a["u0"], g, a["d0a"])
        gs0 = g
        g = r(up_bwd(g))
        g = layer("d1b", a["d1a"], g, a["d1b"])
        g = layer("d1a", a["u1"], g, a["d1a"])
        gs1 = g
        g = r(up_bwd(g))
        g = layer("d2b", a["d2a"], g, a["d2b"])
        g = layer("d2a", a["u2"], g, a["d2a"])
        gs2 = g
        g = r(up_bwd(g))
        g = layer("b1", a["b0"], g, a["b1"])
        g = layer("b0", a["p2"], g, a["b0"])
        g = r(r(pool_bwd(g)) + gs2)
        g = layer("e2b", a["e2a"], g, a["e2b"])
        g = layer("e2a", a["p1"], g, a["e2a"])
        g = r(r(pool_bwd(g)) + gs1)
        g = layer("e1b", a["e1a"], g, a["e1b"])
        g = layer("e1a", a["p0"], g, a["e1a"])
        g = r(r(pool_bwd(g)) + gs0)
        g = layer("e0b", a["e0a"], g, a["e0b"])
        layer("e0a", a["in"], g, a["e0a"], need_dx=False)
        return grads
