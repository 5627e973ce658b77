"""Pins the float64 U-Net oracle (oracle/unet.py) before it is used to check the GPU kernels:
conv3 against brute-force loops, pool/upsample against their adjoints, and the full
backward against central finite differences of the forward."""
import numpy as np
import pytest

from oracle import unet as U


def test_conv3_matches_bruteforce():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 4, 4, 4, 3))
    w = rng.normal(size=(2, 27, 3))
    b = rng.normal(size=2)
    np.testing.assert_allclose(U.conv3(x, w, b), U.conv3_bruteforce(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv3_tap_order():
    """A one-hot weight on tap (dz, dy, dx) must shift the input by +d (y[v] = x[v + d])."""
    rng = np.random.default_rng(4)
    x = rng.normal(size=(1, 4, 4, 4, 1))
    for t, d in enumerate(U.TAPS):
        w = np.zeros((1, 27, 1))
        w[0, t, 0] = 1.0
        y = U.conv3(x, w)
        assert y[0, 1, 2, 3, 0] == x[0, (1 + d[0]) % 4, (2 + d[1]) % 4, (3 + d[2]) % 4, 0]
    assert U.TAPS[13] == (0, 0, 0)


def test_pool_up_adjoint():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(2, 8, 8, 8, 3))
    y = rng.normal(size=(2, 4, 4, 4, 3))
    assert np.isclose(np.sum(U.pool(x) * y), np.sum(x * U.pool_bwd(y)), rtol=1e-12)
    assert np.isclose(np.sum(U.up(y) * x), np.sum(y * U.up_bwd(x)), rtol=1e-12)
    # pool of a constant is the constant; up then pool is the identity
    np.testing.assert_allclose(U.pool(U.up(y)), y, rtol=1e-14)


def test_conv3_bwd_finite_differences():
    rng = np.random.default_rng(6)
    x = rng.normal(size=(1, 4, 4, 4, 3))
    w = rng.normal(size=(2, 27, 3))
    b = rng.normal(size=2)
    r = rng.normal(size=(1, 4, 4, 4, 2))
    f = lambda x, w, b: np.sum(U.conv3(x, w, b) * r)
    dx, dw, db = U.conv3_bwd(x, w, r)
    h = 1e-6
    for idx in [(0, 1, 2, 3, 1), (0, 0, 0, 0, 0), (0, 3, 3, 3, 2)]:
        e = np.zeros_like(x)
        e[idx] = h
        assert np.isclose((f(x + e, w, b) - f(x - e, w, b)) / (2 * h), dx[idx], rtol=1e-7)
    for idx in [(1, 13, 2), (0, 0, 0), (1, 26, 1)]:
        e = np.zeros_like(w)
        e[idx] = h
        assert np.isclose((f(x, w + e, b) - f(x, w - e, b)) / (2 * h), dw[idx], rtol=1e-7)
    assert np.allclose(db, r.sum(axis=(0, 1, 2, 3)))


def test_unet_backward_finite_differences():
    """Whole network at 16^3 (levels 16, 8, 4, 2): analytic gradients of every parameter
    tensor against central differences at two entries each."""
    rng = np.random.default_rng(7)
    p = U.init_params(seed=1)
    x0 = np.zeros((2, 16, 16, 16, 8))
    x0[..., 0] = rng.normal(size=(2, 16, 16, 16))
    x0[..., 1] = rng.normal(size=(2, 16, 16, 16))
    target = rng.normal(size=(2, 16, 16, 16))
    out, a = U.forward(p, x0)
    g = U.backward(p, a, out, target)
    h = 1e-7  # small enough that a ReLU kink is rarely inside [w - h, w + h]
    pick = np.random.default_rng(8)
    for key, val in p.items():
        flat = val.reshape(-1)
        gf = np.abs(g[key].reshape(-1))
        nz = np.flatnonzero(gf)
        assert len(nz) > 0, key  # some unit of every layer is live at this input
        for j in (int(np.argmax(gf)), int(pick.choice(nz))):
            old = flat[j]
            flat[j] = old + h
            lp = U.loss(U.forward(p, x0)[0], target)
            flat[j] = old - h
            lm = U.loss(U.forward(p, x0)[0], target)
            flat[j] = old
            fd = (lp - lm) / (2 * h)
            an = g[key].reshape(-1)[j]
            assert abs(fd - an) <= 1e-5 * abs(an) + 1e-8, (key, j, fd, an)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.normal(size=10000) * 10.0 ** rng.integers(-8, 8, 10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -2.5e-40]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(U.bf16(x), ref)


def test_flatten_roundtrip():
    p = U.init_params(seed=2)
    v = U.flatten(p)
    assert len(v) == sum(27 * ci * co + co for _, ci, co in U.LAYERS) + 17
    q = U.unflatten(v)
    for k in p:
        np.testing.assert_array_equal(p[k], q[k])
