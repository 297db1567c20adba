"""The warm-start U-Net's HIP kernels against the FP64 numpy oracle
(oracle/unet_oracle.py, itself pinned by brute force and finite differences).

The oracle rounds weights, activations and MFMA operands to BF16 at the same
points as the GPU, so the remaining differences come from FP32 vs FP64
accumulation (and the occasional BF16 rounding flip it causes): forward
outputs within 3e-3 and per-layer gradients within 1e-2 relative L2.
"""
import numpy as np
import pytest

import unet_oracle as U

pytestmark = pytest.mark.gpu

LEVELS = 4


def setup_net(seed=7, levels=LEVELS):
    from llr_hip import unet

    rng = np.random.default_rng(seed)
    P = U.init_params(seed, 16, 16, levels)
    P["out"][0] = rng.standard_normal(P["out"][0].shape) * 0.2
    for name in P:
        P[name][1] = rng.standard_normal(P[name][1].shape) * 0.05
    net = unet.UNet(16, 16, levels, seed=1)
    net.set_params(unet.dict_to_params(P, 16, 16, levels))
    return net, P, rng


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_unet_param_layout_round_trip(llr):
    from llr_hip import unet

    net, P, _ = setup_net()
    Q = unet.params_to_dict(net.get_params())
    for name in P:
        assert np.array_equal(Q[name][0], P[name][0].astype(np.float32))
        assert np.array_equal(Q[name][1], P[name][1].astype(np.float32))
    net.close()


# (8, 8, 256): rows of 256 and 128 at levels 0 and 1 take the 3^3 kernel's
# row-aligned path (neighbour validity from the staging, no per-lane masks)
@pytest.mark.parametrize("dims", [(16, 16, 16), (8, 24, 16), (8, 8, 256)])
def test_unet_forward_vs_oracle(llr, dims):
    net, P, rng = setup_net()
    x = rng.standard_normal(dims + (16,)).astype(np.float32)
    y = net.forward(x)
    ref, _ = U.forward(P, x.astype(np.float64), LEVELS)
    e = rel(y - x, ref - x)
    print(f"[unet forward {dims}] rel err of the network part {e:.2e}", flush=True)
    assert e < 3e-3
    net.close()


def gpu_cache(net, x, levels=LEVELS):
    """The oracle's backward cache built from the GPU's own stored activations,
    so both backward passes see the same ReLU masks."""
    D, H, W = x.shape[:3]
    ch = [16 << l for l in range(levels)]
    sh = lambda l, c: (D >> l, H >> l, W >> l, c)
    cache = {"x": U.bf16(x)}
    for l in range(levels):
        a = net.activation("a", l, sh(l, ch[l])).astype(np.float64)
        b = net.activation("b", l, sh(l, ch[l])).astype(np.float64)
        cache[f"enc{l}a.in"] = cache["x"] if l == 0 else net.activation("pool", l, sh(l, ch[l - 1])).astype(np.float64)
        cache[f"enc{l}a.out"], cache[f"enc{l}b.in"], cache[f"enc{l}b.out"] = a, a, b
        if l < levels - 1:
            da = net.activation("da", l, sh(l, ch[l])).astype(np.float64)
            cache[f"dec{l}a.in"] = net.activation("cat", l, sh(l, ch[l] + ch[l + 1])).astype(np.float64)
            cache[f"dec{l}a.out"], cache[f"dec{l}b.in"] = da, da
            cache[f"dec{l}b.out"] = net.activation("db", l, sh(l, ch[l])).astype(np.float64)
    cache["out.in"] = cache["dec0b.out"]
    return cache


@pytest.mark.parametrize("dims", [(16, 16, 16), (24, 32, 48), (8, 16, 128)])
def test_unet_gradients_vs_oracle(llr, dims):
    """Backward kernels against the oracle's backward on the GPU's own forward
    activations (with independent forwards, BF16 rounding flips change ReLU
    masks and the gradients by several percent: the comparison would test the
    chaos, not the kernels). The forward itself is checked above. The larger
    shape splits every weight-gradient reduction over many workgroups."""
    from llr_hip import unet

    net, P, rng = setup_net(9)
    x = rng.standard_normal(dims + (16,)).astype(np.float32)
    gy = rng.standard_normal(dims + (16,)).astype(np.float32)
    g = unet.params_to_dict(net.gradients(x, gy))
    G = U.backward(P, gpu_cache(net, x.astype(np.float64)), gy.astype(np.float64), LEVELS)
    errs = {(name, w): rel(g[name][w], G[name][w]) for name in P for w in (0, 1)}
    print("[unet gradients] per-layer rel err " + " ".join(f"{n}{'wb'[w]}={e:.1e}" for (n, w), e in errs.items()),
          flush=True)
    # FP32-vs-FP64 accumulation flips a few BF16 roundings of the gradients at
    # every layer; the flips accumulate towards the first layers
    assert max(errs.values()) < 5e-3
    net.close()


def test_unet_adam_step_vs_oracle(llr):
    from llr_hip import unet

    net, P, rng = setup_net(11)
    dims = (16, 16, 8)
    x = rng.standard_normal(dims + (16,)).astype(np.float32)
    t = rng.standard_normal(dims + (16,)).astype(np.float32)
    p0 = net.get_params().astype(np.float64)
    pred = net.forward(x).astype(np.float64)
    g = net.gradients(x, (2.0 * (pred - t) / pred.size).astype(np.float32)).astype(np.float64)
    lr = 1e-3
    loss = net.train_step(x, t, lr)
    assert abs(loss - np.mean((pred - t) ** 2)) < 1e-5 * loss
    p1 = net.get_params().astype(np.float64)
    ref = p0.copy()
    U.adam_step(ref, g, np.zeros_like(g), np.zeros_like(g), 1, lr=lr)
    d = np.abs((p1 - p0) - (ref - p0))
    # gradients are summed with atomics (order varies between the two passes):
    # only the signs of vanishing gradients may differ
    assert np.mean(d <= 1e-3 * lr) > 0.999
    assert d.max() <= 2.0 * lr * (1 + 1e-3)
    net.close()


def test_unet_training_reduces_loss(llr):
    from llr_hip import unet

    net = unet.UNet(16, 16, LEVELS, seed=3)
    rng = np.random.default_rng(5)
    dims = (16, 16, 16)
    x = rng.standard_normal(dims + (16,)).astype(np.float32)
    # target: a smooth nonlinear map of the input (needs the network, not the identity)
    t = x + 0.5 * np.tanh(np.roll(x, 1, axis=-1)) * (np.abs(x) > 0.5)
    losses = [net.train_step(x, t, 2e-3) for _ in range(150)]
    print(f"[unet training] loss {losses[0]:.4e} -> {losses[50]:.4e} -> {losses[100]:.4e} -> {losses[-1]:.4e}",
          flush=True)
    assert losses[-1] < 0.7 * losses[0]
    net.close()
