"""CPU pinning of the learned warm start's oracle network (oracle/learned.py)."""
import numpy as np
import numpy.testing as npt

from oracle import learned as L


def _params(seed=0):
    rng = np.random.default_rng(seed)
    return {name: (rng.standard_normal((27, cout, cin)) * np.sqrt(2.0 / (27 * cin)), 0.1 * rng.standard_normal(cout))
            for name, _, cin, cout, _ in L.SPEC}


def test_conv_matches_direct_loops():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((3, 4, 5, 2))
    W = rng.standard_normal((27, 3, 2))
    Y = L.conv_fwd(X, W)
    nz, ny, nx, _ = X.shape
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                ref = np.zeros(3)
                for t, (dx, dy, dz) in enumerate(L.TAPS):
                    if 0 <= x + dx < nx and 0 <= y + dy < ny and 0 <= z + dz < nz:
                        ref += W[t] @ X[z + dz, y + dy, x + dx]
                npt.assert_allclose(Y[z, y, x], ref, rtol=1e-12, atol=1e-12)


def test_adjoints_of_linear_pieces():
    """<A x, y> = <x, A^T y> for the convolution, pooling, upsampling and element-to-node maps."""
    rng = np.random.default_rng(2)
    X = rng.standard_normal((4, 4, 6, 3))
    W = rng.standard_normal((27, 5, 3))
    G = rng.standard_normal((4, 4, 6, 5))
    npt.assert_allclose(np.sum(L.conv_fwd(X, W) * G), np.sum(X * L.conv_bwd_data(G, W)), rtol=1e-11)
    dW, _ = L.conv_bwd_weight(X, G)
    W2 = rng.standard_normal(W.shape)
    npt.assert_allclose(np.sum(L.conv_fwd(X, W2) * G), np.sum(W2 * dW), rtol=1e-11)
    Gp = rng.standard_normal((2, 2, 3, 3))
    npt.assert_allclose(np.sum(L.pool(X) * Gp), np.sum(X * L.pool_bwd(Gp)), rtol=1e-12)
    Gu = rng.standard_normal((8, 8, 12, 3))
    npt.assert_allclose(np.sum(L.up(X) * Gu), np.sum(X * L.up_bwd(Gu)), rtol=1e-12)
    Xe = rng.standard_normal((4, 4, 6, 16))
    dU = rng.standard_normal(5 * 5 * 7 * 3)
    npt.assert_allclose(np.sum(L.elem_to_node(Xe) * dU), np.sum(Xe * L.elem_to_node_bwd(dU, (4, 4, 6), 16)),
                        rtol=1e-12)


def test_network_gradient_vs_finite_differences():
    """Every layer's weight and bias gradient of 0.5*||u - t||^2 against central differences (FP64)."""
    nel = (32, 16, 16)
    rng = np.random.default_rng(3)
    params = _params(4)
    X = rng.standard_normal((nel[2], nel[1], nel[0], 8))
    t = rng.standard_normal(3 * (nel[0] + 1) * (nel[1] + 1) * (nel[2] + 1))
    net = L.Net(params)

    def loss():
        return 0.5 * np.sum((net.forward(X) - t) ** 2)
    u = net.forward(X)
    grads = net.backward(u - t)
    h = 1e-6
    for name, _, cin, cout, _ in L.SPEC:
        W, b = params[name]
        for _ in range(2):
            idx = (rng.integers(27), rng.integers(cout), rng.integers(cin))
            w0 = W[idx]
            W[idx] = w0 + h
            lp = loss()
            W[idx] = w0 - h
            lm = loss()
            W[idx] = w0
            fd = (lp - lm) / (2 * h)
            npt.assert_allclose(grads[name][0][idx], fd, rtol=2e-4, atol=1e-6 * max(1.0, abs(fd)))
        j = rng.integers(cout)
        b0 = b[j]
        b[j] = b0 + h
        lp = loss()
        b[j] = b0 - h
        lm = loss()
        b[j] = b0
        npt.assert_allclose(grads[name][1][j], (lp - lm) / (2 * h), rtol=2e-4, atol=1e-6)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.140625], dtype=np.float64)
    r = L.bf16_round(x)
    npt.assert_array_equal(r, [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 3.140625])
