"""CPU oracle of the warm-start U-Net (test infrastructure only: the library never calls it).

The reference has no learned component; this network is the project's own design, so the
oracle is pinned by mathematics instead of reference vectors:
  * conv3 against a brute-force six-fold loop on tiny inputs (tests/test_unet_oracle.py);
  * every backward against central finite differences of the float64 forward;
  * pool/upsample adjointness <pool x, y> = <x, pool^T y>.

Layout matches the GPU kernels: activations channels-last [B, S, S, S, C] (z, y, x order,
x fastest), 3x3x3 periodic convolutions with weights [Cout, 27, Cin], tap t = 9 (dz+1) +
3 (dy+1) + (dx+1), y[v] = sum_t sum_ci x[v + d_t] w[co, t, ci] + b[co].

Architecture (levels S, S/2, S/4, S/8; ReLU after every 3x3x3 conv):
    e1 = L(x0; 8->16)                       at S     (stored inside c1[..., 32:48])
    e2 = L(L(pool e1; 16->32); 32->32)      at S/2   (inside c2[..., 64:96])
    e3 = L(L(pool e2; 32->64); 64->64)      at S/4   (inside c3[..., 128:192])
    b  = L(L(pool e3; 64->128); 128->128)   at S/8
    d3 = L(L([up b, e3]; 192->64); 64->64)
    d2 = L(L([up d3, e2]; 96->32); 32->32)
    d1 = L([up d2, e1]; 48->16)
    out = d1 . wh + bh                      (1x1 head, no activation)
"""
import numpy as np

LAYERS = [  # name, Cin, Cout
    ("e1", 8, 16), ("e2a", 16, 32), ("e2b", 32, 32), ("e3a", 32, 64), ("e3b", 64, 64),
    ("ba", 64, 128), ("bb", 128, 128), ("d3a", 192, 64), ("d3b", 64, 64),
    ("d2a", 96, 32), ("d2b", 32, 32), ("d1", 48, 16),
]
TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def init_params(seed=0, scale=1.0):
    """He-initialised weights (float64); the GPU side draws its own, tests copy these over."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in LAYERS:
        p[name + ".w"] = rng.normal(0.0, scale * np.sqrt(2.0 / (27 * ci)), (co, 27, ci))
        p[name + ".b"] = rng.normal(0.0, 0.01, co)
    p["head.w"] = rng.normal(0.0, np.sqrt(1.0 / 16), 16)
    p["head.b"] = np.zeros(1)
    return p


def shift(x, d):
    """x[v + d] with periodic wrap (x: [B, S, S, S, C])."""
    return np.roll(x, shift=(-d[0], -d[1], -d[2]), axis=(1, 2, 3))


def conv3(x, w, b=None):
    y = np.zeros(x.shape[:4] + (w.shape[0],))
    for t, d in enumerate(TAPS):
        y += shift(x, d) @ w[:, t, :].T
    if b is not None:
        y += b
    return y


def conv3_bwd(x, w, dy):
    """(dx, dw, db) of y = conv3(x, w, b) for upstream gradient dy."""
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t, d in enumerate(TAPS):
        xs = shift(x, d)
        dw[:, t, :] = np.einsum("bzyxo,bzyxi->oi", dy, xs)
        dx += shift(dy @ w[:, t, :], (-d[0], -d[1], -d[2]))
    return dx, dw, dy.sum(axis=(0, 1, 2, 3))


def pool(x):
    B, S, _, _, C = x.shape
    return x.reshape(B, S // 2, 2, S // 2, 2, S // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(dp):
    return np.repeat(np.repeat(np.repeat(dp, 2, axis=1), 2, axis=2), 2, axis=3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, axis=1), 2, axis=2), 2, axis=3)


def up_bwd(dy):
    B, S, _, _, C = dy.shape
    return dy.reshape(B, S // 2, 2, S // 2, 2, S // 2, 2, C).sum(axis=(2, 4, 6))


def relu(x):
    return np.maximum(x, 0.0)


def ident(x):
    return x


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64: the GPU's storage rounding."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, x0, q=ident):
    """Returns (out [B, S, S, S], cache of activations).  q rounds every stored tensor and the
    convolution weights (bf16 mirrors the GPU kernels; ident is exact float64)."""
    a = {"q": q}

    def L(name, x):
        y = q(relu(conv3(x, q(p[name + ".w"]), p[name + ".b"])))
        a[name + ".in"] = x
        a[name] = y
        return y

    x0 = q(x0)
    e1 = L("e1", x0)
    e2 = L("e2b", L("e2a", q(pool(e1))))
    e3 = L("e3b", L("e3a", q(pool(e2))))
    b = L("bb", L("ba", q(pool(e3))))
    d3 = L("d3b", L("d3a", np.concatenate([up(b), e3], axis=-1)))
    d2 = L("d2b", L("d2a", np.concatenate([up(d3), e2], axis=-1)))
    d1 = L("d1", np.concatenate([up(d2), e1], axis=-1))
    out = d1 @ p["head.w"] + p["head.b"][0]
    return out, a


def loss(out, target):
    """Mean over samples of the voxel-mean squared error."""
    return float(np.mean((out - target) ** 2))


def backward(p, a, out, target):
    """Gradients of loss(out, target) with respect to every parameter.  With the rounding q of
    the forward pass, each gradient tensor the GPU stores is rounded where it is stored: the
    masked input gradient of a layer, and the unmasked data gradients into concatenations and
    pooled tensors."""
    q = a["q"]
    g = {}
    B = out.shape[0]
    nv = out[0].size
    dout = 2.0 * (out - target) / (B * nv)
    d1 = a["d1"]
    g["head.w"] = np.einsum("bzyx,bzyxc->c", dout, d1)
    g["head.b"] = np.array([dout.sum()])
    dy = dout[..., None] * p["head.w"]

    def Lb(name, dy):
        dpre = q(dy * (a[name] > 0))
        dx, g[name + ".w"], g[name + ".b"] = conv3_bwd(a[name + ".in"], q(p[name + ".w"]), dpre)
        return dx

    dc1 = q(Lb("d1", dy))
    dd2, de1 = up_bwd(dc1[..., :32]), dc1[..., 32:]
    dc2 = q(Lb("d2a", Lb("d2b", dd2)))
    dd3, de2 = up_bwd(dc2[..., :64]), dc2[..., 64:]
    dc3 = q(Lb("d3a", Lb("d3b", dd3)))
    db, de3 = up_bwd(dc3[..., :128]), dc3[..., 128:]
    dp3 = q(Lb("ba", Lb("bb", db)))
    dp2 = q(Lb("e3a", Lb("e3b", pool_bwd(dp3) + de3)))
    dp1 = q(Lb("e2a", Lb("e2b", pool_bwd(dp2) + de2)))
    Lb("e1", pool_bwd(dp1) + de1)
    return g


def flatten(p):
    """Parameters (or gradients) in the library's order: per layer W then b, then the head."""
    parts = []
    for name, _, _ in LAYERS:
        parts += [p[name + ".w"].reshape(-1), p[name + ".b"].reshape(-1)]
    parts += [p["head.w"].reshape(-1), p["head.b"].reshape(-1)]
    return np.concatenate(parts)


def unflatten(v):
    p, o = {}, 0
    for name, ci, co in LAYERS:
        p[name + ".w"] = v[o:o + co * 27 * ci].reshape(co, 27, ci)
        o += co * 27 * ci
        p[name + ".b"] = v[o:o + co]
        o += co
    p["head.w"] = v[o:o + 16]
    p["head.b"] = v[o + 16:o + 17]
    assert o + 17 == len(v)
    return p


def conv3_bruteforce(x, w, b):
    """Six nested loops, for pinning conv3 on tiny inputs."""
    B, S, _, _, Ci = x.shape
    Co = w.shape[0]
    y = np.zeros((B, S, S, S, Co))
    for bb in range(B):
        for z in range(S):
            for yy in range(S):
                for xx in range(S):
                    for co in range(Co):
                        s = b[co]
                        for t, (dz, dy, dx) in enumerate(TAPS):
                            s += np.dot(x[bb, (z + dz) % S, (yy + dy) % S, (xx + dx) % S, :], w[co, t, :])
                        y[bb, z, yy, xx, co] = s
    return y
