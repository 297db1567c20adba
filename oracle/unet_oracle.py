"""FP64 numpy restatement of the warm-start 3D U-Net (layers, network,
gradients, Adam) -- the checker for the HIP kernels in unet.hip.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product never
imports this module.

The network has no counterpart in the reference (it is a north-star component
pinned by closed forms): this file is pinned by brute-force convolution loops
and central finite differences (tests/test_unet_oracle.py).

Conventions shared with the HIP code:
  * activations are (D, H, W, C) arrays (channels fastest; x = W fastest of
    the spatial axes), channel counts multiples of 16;
  * a k x k x k convolution (k = 3 with zero padding 1, or k = 1) has weights
    w[o, co, ci] with offset o = (dz*k + dy)*k + dx and bias b[co]:
        y[z, y, x, co] = b[co] + sum_{o, ci} x[z+dz-p, y+dy-p, x+dx-p, ci] w[o, co, ci];
  * BF16 storage is emulated by `bf16` (round to nearest even) wherever the
    GPU stores BF16: weights, activations and the MFMA operands.
"""
import numpy as np


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _offsets(k):
    p = k // 2
    return [(dz - p, dy - p, dx - p) for dz in range(k) for dy in range(k) for dx in range(k)]


def _shift(x, dz, dy, dx):
    """s[z, y, x] = x[z+dz, y+dy, x+dx] with zeros outside."""
    D, H, W = x.shape[:3]
    out = np.zeros_like(x)
    zs, ys, xs = slice(max(0, -dz), min(D, D - dz)), slice(max(0, -dy), min(H, H - dy)), slice(max(0, -dx), min(W, W - dx))
    zt, yt, xt = slice(max(0, dz), min(D, D + dz)), slice(max(0, dy), min(H, H + dy)), slice(max(0, dx), min(W, W + dx))
    out[zs, ys, xs] = x[zt, yt, xt]
    return out


def conv_fwd(x, w, b=None, relu=False):
    k = round(w.shape[0] ** (1 / 3))
    y = np.zeros(x.shape[:3] + (w.shape[1],))
    for o, (dz, dy, dx) in enumerate(_offsets(k)):
        y += _shift(x, dz, dy, dx) @ w[o].T
    if b is not None:
        y += b
    return np.maximum(y, 0.0) if relu else y


def conv_bwd_data(g, w):
    """dL/dx for y = conv(x, w): sum over offsets of shift(g, -o) @ w[o]."""
    k = round(w.shape[0] ** (1 / 3))
    dx = np.zeros(g.shape[:3] + (w.shape[2],))
    for o, (dz, dy, dxx) in enumerate(_offsets(k)):
        dx += _shift(g, -dz, -dy, -dxx) @ w[o]
    return dx


def conv_bwd_weight(g, x, k):
    """dL/dw[o, co, ci] = sum_v g[v, co] x[v + o, ci]; dL/db = sum_v g[v]."""
    G = g.reshape(-1, g.shape[-1])
    dw = np.stack([G.T @ _shift(x, dz, dy, dx).reshape(-1, x.shape[-1]) for dz, dy, dx in _offsets(k)])
    return dw, G.sum(axis=0)


def pool_fwd(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def pool_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up_fwd(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_bwd(g):
    D, H, W, C = g.shape
    return g.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


# --------------------------------------------------------------- network
def layer_specs(cin=16, base=16, levels=4):
    """(name, k, cin, cout, relu) of every convolution in execution order."""
    ch = [base * (2 ** l) for l in range(levels)]
    specs, c = [], cin
    for l in range(levels):
        specs += [(f"enc{l}a", 3, c, ch[l], True), (f"enc{l}b", 3, ch[l], ch[l], True)]
        c = ch[l]
    for l in range(levels - 2, -1, -1):
        specs += [(f"dec{l}a", 3, c + ch[l], ch[l], True), (f"dec{l}b", 3, ch[l], ch[l], True)]
        c = ch[l]
    specs.append(("out", 1, c, cin, False))
    return specs


def init_params(seed, cin=16, base=16, levels=4):
    """He-normal weights (zero output layer: the net starts as the identity map), zero biases."""
    rng = np.random.default_rng(seed)
    P = {}
    for name, k, ci, co, _ in layer_specs(cin, base, levels):
        std = np.sqrt(2.0 / (k ** 3 * ci))
        w = rng.standard_normal((k ** 3, co, ci)) * std
        if name == "out":
            w = np.zeros_like(w)
        P[name] = [w, np.zeros(co)]
    return P


def forward(P, x, levels=4, emulate_bf16=True):
    """Returns (prediction, cache). prediction = x + net(x) (residual), float64.
    With emulate_bf16, weights, stored activations and conv operands are rounded
    to BF16 as on the GPU (gradients stay FP32/FP64)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    cache = {"x": r(x)}
    h = cache["x"]
    skips = []
    for l in range(levels):
        for s in "ab":
            name = f"enc{l}{s}"
            cache[name + ".in"] = h
            h = r(conv_fwd(h, r(P[name][0]), P[name][1], relu=True))
            cache[name + ".out"] = h
        if l < levels - 1:
            skips.append(h)
            h = r(pool_fwd(h))
    for l in range(levels - 2, -1, -1):
        u = r(up_fwd(h))
        h = np.concatenate([u, skips[l]], axis=-1)
        for s in "ab":
            name = f"dec{l}{s}"
            cache[name + ".in"] = h
            h = r(conv_fwd(h, r(P[name][0]), P[name][1], relu=True))
            cache[name + ".out"] = h
    cache["out.in"] = h
    y = conv_fwd(h, r(P["out"][0]), P["out"][1], relu=False)
    return np.asarray(x, np.float64) + y, cache


def backward(P, cache, gy, levels=4, emulate_bf16=True):
    """Gradients of sum(gy * prediction) w.r.t. every weight and bias (dict of [dw, db])."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    G = {}
    base = [P[f"enc{l}a"][0].shape[1] for l in range(levels)]

    def conv_back(name, g, k):
        x = cache[name + ".in"]
        dw, _ = conv_bwd_weight(r(g), x, k)
        G[name] = [dw, g.reshape(-1, g.shape[-1]).sum(axis=0)]
        return conv_bwd_data(r(g), r(P[name][0]))

    g = conv_back("out", gy, 1)
    skip_g = [None] * levels
    # decoder, finest level first (the reverse of the forward order)
    for l in range(0, levels - 1):
        for s in "ba":
            name = f"dec{l}{s}"
            g = g * (cache[name + ".out"] > 0)
            g = conv_back(name, g, 3)
        cu = g.shape[-1] - base[l]
        skip_g[l] = g[..., cu:]
        g = up_bwd(g[..., :cu])
    for l in range(levels - 1, -1, -1):
        if l < levels - 1:
            g = pool_bwd(g) + skip_g[l]
        for s in "ba":
            name = f"enc{l}{s}"
            g = g * (cache[name + ".out"] > 0)
            g = conv_back(name, g, 3)
    return G


def adam_step(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Alg. 1) in place; t counts from 1."""
    m[...] = b1 * m + (1 - b1) * g
    v[...] = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    p[...] = p - lr * mh / (np.sqrt(vh) + eps)
