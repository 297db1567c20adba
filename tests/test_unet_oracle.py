"""Pins the numpy U-Net oracle: brute-force convolution, exact adjoints of the
linear layers, central finite differences of the whole network, and Adam's
closed-form first step."""
import numpy as np
import pytest

import unet_oracle as U


def test_bf16_rounding_kat():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float64)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 rounds up to 1 + 2^-6... (next bf16 above 1+2^-7)
    got = U.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -6 and got[3] == -2.5
    assert abs(got[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8


@pytest.mark.parametrize("k", [1, 3])
def test_conv_matches_brute_force(k):
    rng = np.random.default_rng(1)
    D, H, W, ci, co = 3, 4, 5, 3, 2
    x = rng.standard_normal((D, H, W, ci))
    w = rng.standard_normal((k ** 3, co, ci))
    b = rng.standard_normal(co)
    y = U.conv_fwd(x, w, b)
    p = k // 2
    ref = np.zeros((D, H, W, co))
    for z in range(D):
        for yy in range(H):
            for xx in range(W):
                acc = b.copy()
                for dz in range(k):
                    for dy in range(k):
                        for dx in range(k):
                            zz, y2, x2 = z + dz - p, yy + dy - p, xx + dx - p
                            if 0 <= zz < D and 0 <= y2 < H and 0 <= x2 < W:
                                acc += w[(dz * k + dy) * k + dx] @ x[zz, y2, x2]
                ref[z, yy, xx] = acc
    assert np.abs(y - ref).max() < 1e-12


@pytest.mark.parametrize("k", [1, 3])
def test_conv_gradients_are_exact_adjoints(k):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 3, 5, 4))
    w = rng.standard_normal((k ** 3, 3, 4))
    g = rng.standard_normal((4, 3, 5, 3))
    # <g, conv(x, w)> is bilinear: its x- and w-gradients are the adjoints
    dx = U.conv_bwd_data(g, w)
    dw, db = U.conv_bwd_weight(g, x, k)
    ex, ew = rng.standard_normal(x.shape), rng.standard_normal(w.shape)
    assert abs(np.sum(g * U.conv_fwd(ex, w)) - np.sum(dx * ex)) < 1e-10
    assert abs(np.sum(g * U.conv_fwd(x, ew)) - np.sum(dw * ew)) < 1e-10
    assert np.allclose(db, g.reshape(-1, 3).sum(0))


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 6, 8, 2))
    g = rng.standard_normal((2, 3, 4, 2))
    assert abs(np.sum(U.pool_fwd(x) * g) - np.sum(x * U.pool_bwd(g))) < 1e-12
    assert abs(np.sum(U.up_fwd(g) * x) - np.sum(g * U.up_bwd(x))) < 1e-12
    assert np.allclose(U.pool_fwd(U.up_fwd(g)), g)


def test_network_gradients_match_finite_differences():
    levels = 3
    rng = np.random.default_rng(4)
    P = U.init_params(5, cin=16, base=16, levels=levels)
    P["out"][0] = rng.standard_normal(P["out"][0].shape) * 0.1  # non-zero output layer
    for name in P:
        P[name][1] = rng.standard_normal(P[name][1].shape) * 0.05
    x = rng.standard_normal((8, 8, 8, 16))
    gy = rng.standard_normal((8, 8, 8, 16))
    _, cache = U.forward(P, x, levels, emulate_bf16=False)
    G = U.backward(P, cache, gy, levels, emulate_bf16=False)

    def loss():
        return float(np.sum(gy * U.forward(P, x, levels, emulate_bf16=False)[0]))

    h = 1e-5
    checked = 0
    for name in ["enc0a", "enc1b", "enc2a", "dec1a", "dec0b", "out"]:
        for which in (0, 1):
            arr = P[name][which]
            for _ in range(2):
                idx = tuple(rng.integers(0, s) for s in arr.shape)
                old = arr[idx]
                arr[idx] = old + h
                lp = loss()
                arr[idx] = old - h
                lm = loss()
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                an = G[name][which][idx]
                assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)), (name, which, idx, fd, an)
                checked += 1
    assert checked == 24


def test_adam_first_step_closed_form():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    m, v = np.zeros(3), np.zeros(3)
    U.adam_step(p, g, m, v, 1, lr=1e-2)
    # bias-corrected m/sqrt(v) = g/|g| at t = 1
    assert np.allclose(p, np.array([1.0, -2.0, 0.5]) - 1e-2 * np.sign(g) / (1 + 1e-8 / np.abs(g)))
