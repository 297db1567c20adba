"""Parity at the workloads' own sizes.

  - 256x256x128, 67 volumes, rank-7 subspace, db2 0.005 (the metric's shape):
    the plain FP32 form after one and two iterations against the CPU FP32
    reference path (the oracle's restatement of lib.rs:227-341) on the same
    measurements and block shifts, bar 1e-4 relative NRMSE; the mixed form with
    half storage (the headline form) against the same, bar 5e-3.
  - The learned warm start at that shape on a held-out phantom seed: it must not
    fall back, must stop in fewer iterations than the plain form and reach the
    plain form's NRMSE vs the true image within 1 %.
  - An uneven batch (a 256x256x128x67 member among smaller ones, 13 and 31
    volumes, autotuned block sizes, partial edge blocks): every member bit for
    bit equal to its single reconstruction, one member against the CPU path.

Measurements are built on the device (the library's forward model plus seeded
noise) and copied to the host for the oracle; both sides see the same bytes.
"""
import numpy as np
import pytest

import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c2(llr):
    from llr_hip import phantom

    ds = phantom.device_dataset("c2", _torch(), keep_truth=True, log=_log)
    yield ds
    ds.clear()
    _torch().cuda.empty_cache()


def _recon(ds, n_iter, **kw):
    from llr_hip import recon

    cfg = ds["cfg"]
    r = recon.Recon(ds["vol"], ds["nq"], ds["shifts"][:n_iter], blk=cfg["blk"], llr_mode="svt",
                    lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"], subspace=ds["subspace"],
                    **kw)
    r.set_measurements(device_ptr=ds["y"].data_ptr())
    return r


@pytest.mark.timeout(900)
def test_c2_plain_fp32_and_fp16_vs_cpu_path(c2):
    ds = c2
    cfg = ds["cfg"]
    imgs, hists = {}, {}
    for n in (1, 2):
        r = _recon(ds, n)
        hists[n] = r.run()
        imgs[n] = r.image()
        lam, kc = r.info()
        r.close()
        _log(f"[c2 plain fp32 gpu] {n} iteration(s): residuals {hists[n]}; lambda {lam:.6g}")
    r = _recon(ds, 2, precision="mixed", storage="fp16")
    h16 = r.run()
    img16 = r.image()
    r.close()
    _log(f"[c2 mixed fp16 gpu] 2 iterations: residuals {h16}")
    y = ds["y"].cpu().numpy()
    _log(f"[c2 cpu reference path] setup + 2 iterations on {O.threads()} threads ...")
    x2, h2, lam32, kc32, x1 = O.recon(y, ds["vol"], ds["nq"], cfg["blk"], ds["shifts"][:2], llr_mode=1,
                                      lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"],
                                      subspace=ds["subspace"], snapshot_after=1)
    del y
    e1 = H.nrmse_big(imgs[1], x1)
    e2 = H.nrmse_big(imgs[2], x2)
    e16 = H.nrmse_big(img16, x2)
    _log(f"[c2 parity] NRMSE gpu-vs-cpu32 after 1 it {e1:.3e}, after 2 it {e2:.3e}; mixed fp16 after 2 it {e16:.3e}; "
         f"lambda gpu {lam:.6g} cpu {lam32:.6g}; residuals cpu {h2}")
    assert np.array_equal(kc, kc32)
    assert abs(lam - lam32) <= 1e-5 * lam32
    assert np.allclose(hists[2], h2, rtol=1e-3) and np.allclose(hists[1], h2[:1], rtol=1e-3)
    assert np.allclose(h16, h2, rtol=2e-2)
    assert e1 <= 1e-4 and e2 <= 1e-4
    assert e16 <= 5e-3


@pytest.mark.timeout(600)
def test_c5_learned_fp16_inference_shape(llr):
    """Config 5's inference: 256x256x128x67 from the held-out seed 9001."""
    import os

    from llr_hip import phantom, recon
    from llr_hip import warm_start as W

    torch = _torch()
    weights = os.path.join(os.path.dirname(recon.__file__), "weights", "warm_start_unet.npz")
    preset = dict(phantom.PRESETS["c2"], seeds=(9001, 9002, 9003, 9004))
    ds = phantom.device_dataset(preset, torch, keep_truth=True, log=_log)
    truth = ds["truth"]

    def nrmse(img):
        return float(torch.linalg.vector_norm((img - truth).to(torch.complex128)) /
                     torch.linalg.vector_norm(truth.to(torch.complex128)))

    img = torch.empty_like(ds["y"])
    r = _recon(ds, ds["cfg"]["n_iter"])
    hist = r.run()
    r.image_to_device(img.data_ptr())
    r.close()
    e_plain = nrmse(img)
    stop = float(np.median(hist[-5:]))
    net = W.load_net(weights)
    to_net, from_net = W.tensor_mixing(ds["bvals"], ds["bvecs"], ds["subspace"])
    rl = _recon(ds, ds["cfg"]["n_iter"], precision="mixed", storage="fp16")
    hl, fell_back = rl.run_learned(net, stop, to_net, from_net, momentum=W.LEARNED_MOMENTUM,
                                   min_iter=W.LEARNED_MIN_ITER)
    rl.image_to_device(img.data_ptr())
    rl.close()
    net.close()
    e_learned = nrmse(img)
    _log(f"[c5 learned fp16, seed 9001] {len(hl)} iterations (fallback {fell_back}), NRMSE vs truth {e_learned:.5f} "
         f"vs plain {e_plain:.5f} ({e_learned / e_plain:.4f} x); stop value {stop:.5e}")
    assert not fell_back
    assert W.LEARNED_MIN_ITER <= len(hl) < ds["cfg"]["n_iter"] // 2 and hl[-1] <= stop
    assert e_learned <= 1.01 * e_plain
    del img, ds
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
def test_c4_batch_at_workload_sizes(llr):
    """Seeds 100 (256x256x128x67, R 4), 107 (96x96x64x31, R 8), 112 (96x96x64x13)
    and 103 (192x192x96x13) of the uneven batch, 3 iterations, autotuned."""
    from llr_hip import phantom, recon

    torch = _torch()
    dss = []
    for s in (100, 107, 112, 103):
        m = phantom.batch_member(s)
        m["n_iter"] = 3
        dss.append(phantom.device_dataset(m, torch, log=_log))
    kw = [dict(vol=d["vol"], nq=d["nq"], shifts=d["shifts"], blk=d["cfg"]["blk"],
               lambda_factor=d["cfg"]["lambda_factor"], wavelet_alpha=d["cfg"]["wavelet_alpha"],
               subspace=d["subspace"], partial_blocks=True) for d in dss]
    b = recon.ReconBatch(kw, autotune=True)
    for i, d in enumerate(dss):
        b.set_measurements(i, device_ptr=d["y"].data_ptr())
    hist = b.run()
    tun = [b.tuning(i) for i in range(len(dss))]
    imgs = [b.image(i) for i in range(len(dss))]
    b.close()
    for i, d in enumerate(dss):
        r = recon.Recon(**dict(kw[i], blk=tun[i]["blk"]))
        r.set_measurements(device_ptr=d["y"].data_ptr())
        h = r.run()
        img = r.image()
        r.close()
        same = np.array_equal(img, imgs[i])
        _log(f"[c4 member {i}: {d['vol']}x{d['nq']} R {d['cfg']['R']}] tuning {tun[i]}; batch image == single: "
             f"{same}; residuals {hist[i]} / {h}")
        assert tun[i]["blk"][0] in (6, 8, 10)
        assert same
        # the tuned FFT tiles per workgroup regroup the residual's partial sums
        assert np.allclose(h, hist[i], rtol=1e-12)
    d = dss[1]
    cfg = d["cfg"]
    y = d["y"].cpu().numpy()
    x32, h32, lam32, _ = O.recon(y, d["vol"], d["nq"], tun[1]["blk"], d["shifts"], llr_mode=1,
                                 lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"],
                                 subspace=d["subspace"], partial_blocks=True)
    e = H.nrmse(imgs[1], x32)
    _log(f"[c4 member 1 vs cpu32, blk {tun[1]['blk']} partial] NRMSE {e:.3e}; lambda {tun[1]['lam']:.6g} / "
         f"{lam32:.6g}; residuals {hist[1]} / {h32}")
    assert e <= 1e-4
    assert np.allclose(hist[1], h32, rtol=1e-3)
    del dss
    torch.cuda.empty_cache()
