"""Learned warm start: coefficient packing, the warm-start entry points of the
reconstruction, the cold-start fallback, and what the trained network gains
over the cold start (llr_recon_warm_start / llr_recon_run_learned).

Bars: packing within one BF16 rounding of the numpy restatement; with an
untrained network (its output layer is zero, so it is the identity map) the
warm start is the zero-filled adjoint projected onto the subspace and made
data-consistent, checked against the same steps through torch and the
signal model to FP32 rounding (1e-5 relative, residual 1e-4); a fallback run
is the cold run itself (bit-identical image).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = dict(vol=(32, 32, 16), n_b0=1, dirs="repulsion", n_dirs=12, bval=3000.0, R=4.0, n_iter=8, blk=(8, 8, 8),
             lambda_factor=0.02, wavelet_alpha=0.005, srank=7, n_ell=8)


def nrmse_t(a, b):
    import torch

    return float(torch.linalg.vector_norm((a - b).to(torch.complex128)) /
                 torch.linalg.vector_norm(b.to(torch.complex128)))


def small_ds(seed=41):
    from llr_hip import warm_start as W

    return W.dataset_device(W.member_config(seed, SMALL))


def test_pack_and_identity_prediction(llr):
    import torch

    from llr_hip import unet

    ds = small_ds()
    vol, K = ds["vol"], 7
    N = int(np.prod(vol))
    rng = np.random.default_rng(3)
    c = (rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))).astype(np.complex64)
    cd = torch.from_numpy(c.ravel()).cuda()
    scale = unet.coefficient_scale_device(cd.data_ptr(), K, vol, ds["to_net"])
    beta = ds["to_net"].astype(np.float64) @ c.astype(np.complex128)  # (K, N)
    assert abs(scale - np.sqrt(np.mean(np.abs(beta[0]) ** 2))) < 1e-5 * scale
    out = torch.empty((vol[2], vol[1], vol[0], 16), dtype=torch.float32, device="cuda")
    unet.pack_coefficients_device(cd.data_ptr(), K, vol, 16, scale, out.data_ptr(), ds["to_net"], False)
    ref = np.zeros((N, 16))
    ref[:, 0:2 * K:2] = (beta.real / scale).T
    ref[:, 1:2 * K:2] = (beta.imag / scale).T
    got = out.cpu().numpy().reshape(N, 16)
    assert np.max(np.abs(got - ref)) < 1e-5 * np.max(np.abs(ref))
    outb = torch.empty((vol[2], vol[1], vol[0], 16), dtype=torch.int16, device="cuda")
    unet.pack_coefficients_device(cd.data_ptr(), K, vol, 16, scale, outb.data_ptr(), ds["to_net"], True)
    gb = (outb.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32).reshape(N, 16)
    assert np.all(np.abs(gb - ref) <= 2.0 ** -8 * np.abs(ref) + 1e-30)
    # a fresh network is the identity map: prediction == input (FP32 rounding of the two mixings)
    net = unet.UNet(16, 16, 4, seed=5)
    pd = torch.empty_like(cd)
    net.predict_coefficients_device(cd.data_ptr(), K, vol, scale, pd.data_ptr(), ds["to_net"], ds["from_net"])
    e = nrmse_t(pd, cd)
    print(f"[identity prediction] rel err {e:.2e}", flush=True)
    assert e < 1e-5
    net.close()


def test_tensor_mixing_is_an_inverse_pair(llr):
    from llr_hip import warm_start as W

    ds = small_ds(43)
    assert np.allclose(ds["to_net"].astype(np.float64) @ ds["from_net"], np.eye(7), atol=1e-5)


def test_coefficients_round_trip(llr):
    import torch

    ds = small_ds(44)
    from llr_hip import warm_start as W

    r = W.recon_for(ds)
    r.prepare()
    K, N = 7, int(np.prod(ds["vol"]))
    c0 = torch.empty(K * N, dtype=torch.complex64, device="cuda")
    r.coefficients_to_device(c0.data_ptr(), c0.numel())
    r.set_coefficients(device_ptr=c0.data_ptr(), n=c0.numel())
    c1 = torch.empty_like(c0)
    r.coefficients_to_device(c1.data_ptr(), c1.numel())
    assert nrmse_t(c1, c0) < 1e-6
    r.close()


def test_identity_warm_start_is_the_consistent_subspace_projection(llr):
    """Untrained (identity) network: the warm start is the zero-filled
    adjoint projected onto the subspace, then made data-consistent. Checked
    against the same steps done through the other entry points and torch:
    the prediction's residual, the data consistency of the result, and the
    threshold (taken from the zero-filled adjoint, as for the cold start)."""
    import torch

    from llr_hip import signal_model, unet
    from llr_hip import warm_start as W

    ds = small_ds(45)
    vol, nq, K = ds["vol"], ds["nq"], 7
    N = int(np.prod(vol))
    cold = W.recon_for(ds)
    h_cold = cold.run()
    img_cold = torch.empty_like(ds["y"])
    cold.image_to_device(img_cold.data_ptr())
    lam_cold, _ = cold.info()
    cold.close()
    # the same projection by hand: P = C S^H of the adjoint's coefficients, then its forward model
    ref = W.recon_for(ds)
    ref.prepare()
    _, kc = ref.info()  # the k-space corrections the reconstruction estimated
    c = torch.empty(K * N, dtype=torch.complex64, device="cuda")
    ref.coefficients_to_device(c.data_ptr(), c.numel())
    S = torch.from_numpy(np.asarray(ds["subspace"])).cuda()  # nq x K
    P = (S.conj() @ c.view(K, N)).reshape(-1).contiguous()  # X[q] = sum_k C[k] conj(S[q][k])
    kp = P.clone()
    torch.cuda.synchronize()
    z = np.zeros((nq, 3), np.float32)
    signal_model.signal_model_device(kp.data_ptr(), kp.numel(), vol, nq, kc, z, signal_model.ModelDirection.Forward)
    sampled = ds["y"] != 0
    energy = float(torch.sum(torch.abs(ds["y"].to(torch.complex128)) ** 2))
    res_ref = float(torch.sum(torch.abs((kp - ds["y"])[sampled].to(torch.complex128)) ** 2)) / energy
    ref.close()

    net = unet.UNet(16, 16, 4, seed=1)
    warm = W.recon_for(ds)
    scale, res0 = warm.warm_start(net, ds["to_net"], ds["from_net"])
    assert scale > 0
    print(f"[identity warm start] prediction residual {res0:.6e} (by hand {res_ref:.6e})", flush=True)
    assert abs(res0 - res_ref) < 1e-4 * res_ref
    img = torch.empty_like(ds["y"])
    warm.image_to_device(img.data_ptr())
    k = img.clone()
    torch.cuda.synchronize()
    signal_model.signal_model_device(k.data_ptr(), k.numel(), vol, nq, kc, z, signal_model.ModelDirection.Forward)
    # data-consistent: sampled k-space equals the measurements, the rest is the projection's
    assert nrmse_t(k[sampled], ds["y"][sampled]) < 1e-5
    assert nrmse_t(k[~sampled], kp[~sampled]) < 1e-5
    h_warm = warm.run()
    lam_warm, _ = warm.info()
    assert lam_warm == lam_cold
    assert len(h_warm) == len(h_cold)
    # learned run with an unreachable stop: falls back to the cold run itself
    hist, fell_back = warm.run_learned(net, 1e-12, ds["to_net"], ds["from_net"])
    assert fell_back and len(hist) == ds["cfg"]["n_iter"]
    warm.image_to_device(img.data_ptr())
    assert torch.equal(img, img_cold)
    assert np.array_equal(hist, h_cold)
    # reachable stop: the warm-started run ends as soon as it reaches it
    thr = float(np.max(h_warm))
    hist, fell_back = warm.run_learned(net, thr, ds["to_net"], ds["from_net"])
    assert not fell_back and hist[-1] <= thr and np.all(hist[:-1] > thr)
    assert len(hist) <= ds["cfg"]["n_iter"]
    warm.close()
    net.close()


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "llr-reco_amd", "llr_hip",
                                                    "weights", "warm_start_unet.npz")),
                    reason="no trained warm-start weights in the tree")
def test_trained_warm_start_gains_over_cold_start(llr):
    """Held-out seed of the training workload (128x128x96, 31 volumes): the
    trained network's start reaches the plain run's final residual in fewer
    iterations than the plain run, at an NRMSE vs the true image no worse than
    1% above the plain run's, and is better than the cold start after the same
    number of iterations."""
    import torch

    from llr_hip import warm_start as W

    net = W.load_net()
    ds = W.dataset_device(W.member_config(9001))
    truth = ds["truth"]
    img = torch.empty_like(ds["y"])
    cold = W.recon_for(ds)
    h_cold = cold.run()
    cold.image_to_device(img.data_ptr())
    e_cold = nrmse_t(img, truth)
    cold.close()
    learned = W.recon_for(ds, precision="mixed")
    hist, fell_back = learned.run_learned(net, float(h_cold[-1]), ds["to_net"], ds["from_net"])
    learned.image_to_device(img.data_ptr())
    e_learned = nrmse_t(img, truth)
    learned.close()
    k = len(hist)
    same = W.recon_for(ds, shifts=ds["shifts"][:k])
    same.run()
    same.image_to_device(img.data_ptr())
    e_same = nrmse_t(img, truth)
    same.close()
    print(f"[trained warm start] iterations {k} vs {len(h_cold)}; NRMSE vs truth learned {e_learned:.5f}, "
          f"cold@{len(h_cold)} {e_cold:.5f}, cold@{k} {e_same:.5f}; fallback {fell_back}", flush=True)
    assert not fell_back
    assert k < len(h_cold)
    assert e_learned <= 1.01 * e_cold
    assert e_learned < e_same
    net.close()


def test_learned_momentum_matches_cpu_path(llr):
    """Momentum of the learned run (coefficient space: (1 + b) C(x_k) - b C(x_{k-1})
    entering the LLR step) against the CPU FP32 reference path's image-space
    form x_k + b (x_k - x_{k-1}) from the same warm-start iterate, three
    iterations, bar 1e-4 relative NRMSE; the same run without momentum matches
    the plain CPU loop from that start, and the two differ by far more than
    the bar."""
    import oracle as O
    import torch

    from llr_hip import unet
    from llr_hip import warm_start as W

    ds = small_ds(46)
    cfg = ds["cfg"]
    net = unet.UNet(16, 16, 4, seed=2)
    y = ds["y"].cpu().numpy()
    img = torch.empty_like(ds["y"])
    r = W.recon_for(ds)
    r.warm_start(net, ds["to_net"], ds["from_net"])
    r.image_to_device(img.data_ptr())
    x0 = img.cpu().numpy()
    lam, _ = r.info()
    got = {}
    for beta in (0.0, 0.6):
        # an always-reached stop after exactly three iterations: no fallback
        h, fell_back = r.run_learned(net, 1e30, ds["to_net"], ds["from_net"], max_iter=3, momentum=beta, min_iter=3)
        assert not fell_back and len(h) == 3
        r.image_to_device(img.data_ptr())
        got[beta] = (img.cpu().numpy(), h)
    r.close()
    net.close()
    for beta in (0.0, 0.6):
        x32, h32, _, _ = O.recon(y, ds["vol"], ds["nq"], cfg["blk"], ds["shifts"][:3], llr_mode=1, svt_lambda=lam,
                                 wavelet_alpha=cfg["wavelet_alpha"], subspace=ds["subspace"], x0=x0, momentum=beta)
        e = float(np.linalg.norm(got[beta][0] - x32) / np.linalg.norm(x32))
        print(f"[learned momentum {beta}] NRMSE vs cpu32 {e:.3e}; residuals {got[beta][1]} / {h32}", flush=True)
        assert e <= 1e-4
        assert np.allclose(got[beta][1], h32, rtol=1e-3)
    d = float(np.linalg.norm(got[0.6][0] - got[0.0][0]) / np.linalg.norm(got[0.0][0]))
    print(f"[learned momentum] 0.6 vs 0 differ by {d:.3e}", flush=True)
    assert d > 1e-3
