"""Warm-start network kernels on the GPU against the CPU oracle (oracle/unet.py).

The oracle runs the same network in float64, once with every stored tensor rounded to
bfloat16 where the GPU stores it (so only FP32 accumulation order differs) and once exactly
(the whole BF16 error).  Sizes: 16^3 samples (levels 16, 8, 4, 2), batch 2."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import unet as U

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def inputs(S, B, seed):
    rng = np.random.default_rng(seed)
    V = -0.5 * np.abs(rng.normal(size=(B, S, S, S)))
    T = rng.normal(size=(B, S, S, S))
    return np.ascontiguousarray(V), np.ascontiguousarray(T)


@pytest.fixture
def net(lg):
    s = lg.grid_surrogate_alloc(16, 2, 2.0, 0)
    yield s
    lg.grid_surrogate_free(s)


def test_params_roundtrip(lg, net):
    case("surrogate params roundtrip")
    flat = U.flatten(U.init_params(seed=3))
    assert lg.grid_surrogate_nparams(net) == len(flat)
    lg.grid_surrogate_set_params(net, dp(flat))
    back = np.zeros_like(flat)
    lg.grid_surrogate_get_params(net, dp(back))
    np.testing.assert_array_equal(back, flat.astype(np.float32).astype(np.float64))


def test_forward_and_gradients_vs_oracle(lg, net):
    case("surrogate forward/backward 16^3 x 2 vs oracle")
    S, B = 16, 2
    flat = U.flatten(U.init_params(seed=3))
    lg.grid_surrogate_set_params(net, dp(flat))
    V, T = inputs(S, B, 11)
    grad = np.zeros_like(flat)
    loss_gpu = lg.grid_surrogate_gradient(net, B, dp(V), dp(T), dp(grad))
    out_gpu = np.zeros((B, S, S, S))
    lg.grid_surrogate_predict(net, B, dp(V), dp(out_gpu))
    p32 = U.unflatten(flat.astype(np.float32).astype(np.float64))
    x0 = np.zeros((B, S, S, S, 8))
    x0[..., 0] = 2.0 * V.astype(np.float32)
    T32 = T.astype(np.float32).astype(np.float64)
    res = {}
    for label, q in (("bf16", U.bf16), ("exact", U.ident)):
        out, a = U.forward(p32, x0, q=q)
        g = U.flatten(U.backward(p32, a, out, T32))
        res[label] = (out, U.loss(out, T32), g)
    out_b, loss_b, g_b = res["bf16"]
    out_e, loss_e, g_e = res["exact"]
    print(f"forward rel err vs bf16 oracle {rel(out_gpu, out_b):.2e}, vs exact {rel(out_gpu, out_e):.2e}; "
          f"loss {loss_gpu:.6f} / {loss_b:.6f} / {loss_e:.6f}")
    assert rel(out_gpu, out_b) < 2e-3
    assert rel(out_gpu, out_e) < 3e-2
    assert abs(loss_gpu - loss_b) < 1e-3 * loss_b
    # per tensor
    gp = U.unflatten(grad)
    gb = U.unflatten(g_b)
    ge = U.unflatten(g_e)
    worst_b = worst_e = 0.0
    for k in gp:
        eb, ee = rel(gp[k], gb[k]), rel(gp[k], ge[k])
        print(f"  {k:8s} |g| {np.linalg.norm(gb[k]):.3e}  rel err vs bf16 oracle {eb:.2e}  vs exact {ee:.2e}")
        worst_b, worst_e = max(worst_b, eb), max(worst_e, ee)
    assert worst_b < 2e-2
    assert worst_e < 1e-1
    assert rel(grad, g_e) < 3e-2


def test_training_reduces_loss(lg):
    """A few epochs of Adam on a smooth regression (target = normalised exp(-2 box-average V))."""
    case("surrogate training smoke 16^3")
    S, n = 16, 16
    rng = np.random.default_rng(5)
    V = -0.5 * np.abs(rng.normal(size=(n, S, S, S)))
    T = np.exp(-2.0 * V)
    T /= np.sqrt(np.mean(T * T, axis=(1, 2, 3), keepdims=True))
    s = lg.grid_surrogate_alloc(S, 4, 2.0, 1)
    try:
        l0 = lg.grid_surrogate_loss(s, n, dp(V), dp(T))
        lg.grid_surrogate_set_training_data(s, n, dp(V), dp(T))
        hist = np.zeros(60)
        lg.grid_surrogate_train(s, 60, 1e-3, 0, dp(hist))
        l1 = lg.grid_surrogate_loss(s, n, dp(V), dp(T))
        print(f"loss {l0:.4f} -> {l1:.4f}; per epoch {np.array2string(hist[::6], precision=4)}")
        assert l1 < 0.2 * l0
    finally:
        lg.grid_surrogate_free(s)


def test_symgather_matches_numpy(lg):
    """The training augmentation: one S^3 sample under an axis permutation, reflections
    i -> -i mod S and a periodic shift, against the same map built with numpy indices."""
    case("symgather 8^3")
    S, lgS = 8, 3
    rng = np.random.default_rng(4)
    src = rng.normal(size=(S, S, S)).astype(np.float32)
    nbytes = src.nbytes
    d_src, d_dst = C.c_void_p(), C.c_void_p()
    assert lg.hgrid_malloc(C.byref(d_src), nbytes) == 0 and lg.hgrid_malloc(C.byref(d_dst), nbytes) == 0
    try:
        assert lg.hgrid_memcpy_h2d(d_src, src.ctypes.data, nbytes) == 0
        P = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
        c = np.indices((S, S, S))
        for perm, flips, sh in [(0, 0, (0, 0, 0)), (3, 5, (1, 7, 2)), (5, 2, (4, 0, 3)), (1, 7, (6, 6, 6))]:
            assert lg.hgrid_nn_symgather(d_src, d_dst, lgS, perm, flips, *sh) == 0
            out = np.empty_like(src)
            assert lg.hgrid_memcpy_d2h(out.ctypes.data, d_dst, nbytes) == 0
            q = []
            for a in range(3):
                v = c[P[perm][a]]
                if (flips >> a) & 1:
                    v = (S - v) % S
                q.append((v + sh[a]) % S)
            np.testing.assert_array_equal(out, src[q[0], q[1], q[2]])
    finally:
        lg.hgrid_free(d_src)
        lg.hgrid_free(d_dst)
