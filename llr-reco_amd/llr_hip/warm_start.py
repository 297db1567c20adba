"""Learned warm start of the LLR reconstruction (no counterpart in the reference).

A 3D U-Net (llr_unet_*, BF16 MFMA forward/backward kernels, Adam) maps the
zero-filled adjoint's subspace coefficients to the coefficients the plain
method converges to; the reconstruction then iterates from that prediction
until its residual reaches the plain method's stopping value, with the plain
cold start as the fallback (llr_recon_run_learned).

Channel space. The rank-K subspace S (nq x K) comes from an SVD of the
log-linear tensor design A (phantom.loglinear_subspace), so its columns depend
on the direction set: a network trained on one b-table would see rotated
channels on another. The network therefore works on the design coefficients
beta = A^+ S c (A with b in units of 1000 s/mm^2): one intercept and six
tensor terms, the same physical quantities for every b-table. to_net = A^+ S
and from_net = S^T A are inverse K x K maps (S S^T A = A). Per data set the
channels are divided by the RMS of beta_0 over the volume.

Training pairs come from the plain method itself: seeded phantoms (built on
the GPU), their measurements, the zero-filled adjoint's coefficients (input)
and the coefficients after the plain run (target), both packed to BF16
network tensors that stay in HBM. Training takes random crops of the stored
pairs.

  python -m llr_hip.warm_start train --pairs 2000 --epochs 20 --out weights.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import phantom, recon, signal_model, unet

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warm_start_unet.npz")

# the training workload: 128x128x96 phantoms with 31 volumes (1 b0 + 30
# directions at b=3000), rank-7 subspace, R=6, 30 plain iterations
TRAIN_MEMBER = dict(vol=(128, 128, 96), n_b0=1, dirs="repulsion", n_dirs=30, bval=3000.0, R=6.0, n_iter=30,
                    blk=(8, 8, 8), lambda_factor=0.02, wavelet_alpha=0.005, srank=7, n_ell=20)
CHANNELS = 16

# Iterations after the warm start (llr_recon_run_learned): momentum b of the
# extrapolation x_k + b (x_k - x_{k-1}) entering the LLR step, and the number
# of iterations before the residual stop is tested. The residual of these runs
# reaches the plain run's plateau in a few iterations while the image error
# still falls for several more, so the stop test alone can come too early.
# Chosen on 6 held-out 256x256x128x67 phantoms (seeds 9001, 9005, 9009, 31, 700,
# 4242): 7-11 iterations (mean 8.0), NRMSE vs the true image at most 1.0054 x
# the plain 30-iteration run's; without momentum and floor, 15-26 (mean 19.5).
LEARNED_MOMENTUM = 0.7
LEARNED_MIN_ITER = 7


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tensor_mixing(bvals, bvecs, subspace, b_unit=1000.0):
    """(to_net, from_net): K x K float32 maps between subspace coefficients and
    log-linear design coefficients (identity when the subspace is not the
    full design space)."""
    S = np.real(np.asarray(subspace, np.complex128))
    A = phantom.loglinear_design(np.asarray(bvals, np.float64) / b_unit, bvecs)
    K = S.shape[1]
    if A.shape[1] != K or np.linalg.matrix_rank(A) < K:
        eye = np.eye(K, dtype=np.float32)
        return eye, eye.copy()
    to_net = np.linalg.pinv(A) @ S
    from_net = S.T @ A
    return to_net.astype(np.float32), from_net.astype(np.float32)


def member_config(seed, base=None):
    cfg = dict(base or TRAIN_MEMBER)
    cfg["seeds"] = (seed, seed + 1000, seed + 2000, seed + 3000)
    return cfg


def phantom_image_device(vol, bvals, bvecs, seed, n_ell=20, device="cuda"):
    """phantom.phantom_image built on the GPU with torch (the same ellipsoids,
    tissue values and phase): flat complex64, x fastest, volume slowest."""
    import torch

    params = phantom.phantom_params(seed, n_ell)
    nx, ny, nz = vol
    u = [(torch.arange(n, dtype=torch.float64, device=device) - n / 2) / (n / 2) for n in (nx, ny, nz)]
    Z, Y, X = torch.meshgrid(u[2], u[1], u[0], indexing="ij")  # (nz, ny, nx): x fastest
    lab = torch.zeros((nz, ny, nx), dtype=torch.uint8, device=device)
    head = (X / 0.85) ** 2 + (Y / 0.9) ** 2 + (Z / 0.8) ** 2 <= 1.0
    lab[head] = 1
    for k, e in enumerate(params["ells"]):
        R, c, ax = e["R"], e["c"], e["ax"]
        dx, dy, dz = X - c[0], Y - c[1], Z - c[2]
        s = 0.0
        for j in range(3):
            q = dx * R[0, j] + dy * R[1, j] + dz * R[2, j]
            s = s + (q / ax[j]) ** 2
        lab[(s <= 1.0) & head] = k + 2
    p = params["phase"]
    ph = p["c0"] + p["lin"][0] * X + p["lin"][1] * Y + p["lin"][2] * Z
    qd = p["quad"]
    ph = ph + qd[0] * X * X + qd[1] * Y * Y + qd[2] * Z * Z + qd[3] * X * Y + qd[4] * X * Z + qd[5] * Y * Z
    phase = torch.polar(torch.ones_like(ph), ph).to(torch.complex64)
    del X, Y, Z, ph
    D_list = [np.zeros((3, 3)), params["bg_D"]] + [e["D"] for e in params["ells"]]
    s0_list = [0.0, params["bg_s0"]] + [e["s0"] for e in params["ells"]]
    amp = np.array([[s0 * np.exp(-b * g @ D @ g) for s0, D in zip(s0_list, D_list)]
                    for b, g in zip(bvals, bvecs)], np.float32)  # (nq, labels)
    amp_t = torch.from_numpy(amp).to(device)
    labl = lab.reshape(-1).long()
    out = amp_t[:, labl].to(torch.complex64) * phase.reshape(1, -1)
    obj = labl > 0
    return out.reshape(-1), obj


def measurements_device(img, obj, vol, nq, R, s_noise, s_mask, ctx=None):
    """Forward model (library, zero corrections), complex Gaussian noise at SNR
    30 of the b0 object mean, variable-density PE mask; device tensor out."""
    import torch

    N = int(np.prod(vol))
    s0_obj = float(torch.mean(torch.abs(img[:N][obj])))
    y = img.clone()
    torch.cuda.synchronize()
    zero = np.zeros((nq, 3), np.float32)
    signal_model.signal_model_device(y.data_ptr(), y.numel(), vol, nq, zero, zero,
                                     signal_model.ModelDirection.Forward, ctx)
    g = torch.Generator(device=y.device)
    g.manual_seed(int(s_noise))
    noise = torch.randn(2 * y.numel(), device=y.device, dtype=torch.float32, generator=g) * (s0_obj / 30.0 / np.sqrt(2.0))
    y += torch.view_as_complex(noise.view(-1, 2))
    del noise
    mask = phantom.pe_mask(vol[1], vol[2], nq, R, s_mask)  # (ny, nz, nq)
    m = torch.from_numpy(np.ascontiguousarray(np.real(mask).transpose(2, 1, 0)) != 0).to(y.device)  # (nq, nz, ny)
    y.view(nq, vol[2], vol[1], vol[0]).mul_(m[..., None].to(torch.complex64))
    torch.cuda.synchronize()
    return y


def dataset_device(cfg, device="cuda", ctx=None):
    """One seeded data set on the GPU: truth image, measurements, b-table,
    subspace, mixing matrices and shift sequence."""
    s_ph, s_noise, s_mask, s_shift = cfg["seeds"]
    dirs = phantom.six_directions() if cfg["dirs"] == "six" else phantom.repulsion_directions(cfg["n_dirs"], s_ph + 1000)
    bvals, bvecs = phantom.b_table(cfg["n_b0"], dirs, cfg["bval"])
    nq = len(bvals)
    vol = tuple(cfg["vol"])
    img, obj = phantom_image_device(vol, bvals, bvecs, s_ph, cfg["n_ell"], device)
    y = measurements_device(img, obj, vol, nq, cfg["R"], s_noise, s_mask, ctx)
    sub = phantom.loglinear_subspace(bvals, bvecs, cfg["srank"])
    to_net, from_net = tensor_mixing(bvals, bvecs, sub)
    shifts = phantom.shift_sequence(vol, cfg["n_iter"], s_shift)
    return dict(truth=img, y=y, vol=vol, nq=nq, bvals=bvals, bvecs=bvecs, subspace=sub, to_net=to_net,
                from_net=from_net, shifts=shifts, cfg=cfg)


def recon_for(ds, precision="fp32", residual_stop=0.0, shifts=None, ctx=None):
    cfg = ds["cfg"]
    r = recon.Recon(ds["vol"], ds["nq"], ds["shifts"] if shifts is None else shifts, blk=cfg["blk"],
                    lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"], subspace=ds["subspace"],
                    precision=precision, residual_stop=residual_stop, ctx=ctx)
    r.set_measurements(device_ptr=ds["y"].data_ptr())
    return r


def make_pair(ds, ctx=None):
    """(input, target) BF16 network tensors (nz, ny, nx, CHANNELS) on the GPU:
    the zero-filled adjoint's coefficients and the plain run's, mixed into
    design space and divided by the adjoint's beta_0 RMS. Also returns the plain
    run's residual history."""
    import torch

    vol, K = ds["vol"], ds["subspace"].shape[1]
    N = int(np.prod(vol))
    r = recon_for(ds, ctx=ctx)
    coef = torch.empty(K * N, dtype=torch.complex64, device=ds["y"].device)
    r.prepare()
    r.coefficients_to_device(coef.data_ptr(), coef.numel())
    scale = unet.coefficient_scale_device(coef.data_ptr(), K, vol, ds["to_net"], ctx)
    shape = (vol[2], vol[1], vol[0], CHANNELS)
    x = torch.empty(shape, dtype=torch.int16, device=coef.device)
    t = torch.empty(shape, dtype=torch.int16, device=coef.device)
    unet.pack_coefficients_device(coef.data_ptr(), K, vol, CHANNELS, scale, x.data_ptr(), ds["to_net"], True, ctx)
    hist = r.run()
    r.coefficients_to_device(coef.data_ptr(), coef.numel())
    unet.pack_coefficients_device(coef.data_ptr(), K, vol, CHANNELS, scale, t.data_ptr(), ds["to_net"], True, ctx)
    r.close()
    return x, t, dict(scale=scale, residual=hist)


def generate_pairs(seeds, base=None, device="cuda", ctx=None, progress=50):
    import torch

    xs, ts, infos = [], [], []
    t0 = time.time()
    for i, s in enumerate(seeds):
        ds = dataset_device(member_config(s, base), device, ctx)
        x, t, info = make_pair(ds, ctx)
        xs.append(x)
        ts.append(t)
        infos.append(info)
        del ds
        if progress and (i + 1) % progress == 0:
            log(f"pairs: {i + 1}/{len(seeds)} in {time.time() - t0:.1f} s "
                f"(HBM in use {torch.cuda.memory_allocated() / 2**30:.1f} GiB)")
    return xs, ts, infos


def crop_origins(rng, full, crop):
    return tuple(int(rng.integers(0, f - c + 1)) for f, c in zip(full, crop))


def bf16_to_f32(a):
    """BF16 bit patterns (int16 tensor) -> float32 tensor."""
    import torch

    return (a.to(torch.int32).bitwise_and(0xFFFF) << 16).view(torch.float32)


def validation_loss(net, vxs, vts):
    """(network, identity map) mean squared error on whole held-out pairs."""
    import torch

    tot = ident = 0.0
    for x, t in zip(vxs, vts):
        xf, tf = bf16_to_f32(x).contiguous(), bf16_to_f32(t)
        y = torch.empty_like(xf)
        torch.cuda.synchronize()
        net.forward_device(xf.data_ptr(), y.data_ptr(), xf.shape)
        torch.cuda.synchronize()
        tot += float(torch.mean((y - tf) ** 2))
        ident += float(torch.mean((xf - tf) ** 2))
    return tot / max(len(vxs), 1), ident / max(len(vxs), 1)


def train(net, xs, ts, epochs=20, lr=1e-3, crop=(48, 64, 64), seed=0, progress=True, val=None, time_budget=None,
          on_epoch=None, warmup=0):
    """Adam on random (z, y, x) crops of the stored pairs, one crop per pair per
    epoch in a shuffled order; the step size ramps linearly to `lr` over the
    first `warmup` steps. Returns per-epoch records (mean training loss,
    validation and identity-map losses when `val` = (xs, ts) is given). Stops
    early, after a whole epoch, once `time_budget` seconds have passed."""
    import torch

    rng = np.random.default_rng(seed)
    full = tuple(xs[0].shape[:3])
    crop = tuple(min(c, f) for c, f in zip(crop, full))
    torch.cuda.synchronize()
    records = []
    t0 = time.time()
    step = 0
    for ep in range(epochs):
        tot = 0.0
        for i in rng.permutation(len(xs)):
            o = crop_origins(rng, full, crop)
            step += 1
            lr_s = lr * min(1.0, step / warmup) if warmup else lr
            tot += net.train_step_crop(xs[i].data_ptr(), ts[i].data_ptr(), full, o, crop, lr_s)
        rec = dict(epoch=ep + 1, train=tot / len(xs), t=round(time.time() - t0, 1))
        if val is not None:
            rec["val"], rec["identity"] = validation_loss(net, *val)
        records.append(rec)
        if progress:
            log("epoch " + json.dumps(rec))
        if on_epoch is not None:
            on_epoch(records)
        if time_budget is not None and time.time() - t0 > time_budget:
            break
    return records


def load_net(path=None, ctx=None):
    return unet.UNet.load(path or WEIGHTS, ctx)


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the warm-start U-Net on pairs from the plain method")
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--first-seed", type=int, default=1000)
    ap.add_argument("--val", type=int, default=8, help="held-out pairs (seeds 5000+) for the validation loss")
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--warmup", type=int, default=0, help="steps of linear learning-rate warm-up")
    ap.add_argument("--crop", type=int, nargs=3, default=[48, 64, 64], help="crop (z y x)")
    ap.add_argument("--time-budget", type=float, default=None, help="stop after the epoch that passes this (s)")
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args(argv)
    import torch

    from . import load

    load()
    t0 = time.time()
    seeds = list(range(args.first_seed, args.first_seed + args.pairs))
    xs, ts, infos = generate_pairs(seeds)
    vx, vt, _ = generate_pairs(list(range(5000, 5000 + args.val)), progress=0)
    t_gen = time.time() - t0
    log(f"generated {len(xs)} + {len(vx)} pairs in {t_gen:.1f} s")
    net = unet.UNet(CHANNELS, 16, 4, seed=args.first_seed)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    meta = dict(pairs=len(xs), seeds=f"{seeds[0]}-{seeds[-1]}", val_seeds=f"5000-{5000 + args.val - 1}",
                lr=args.lr, warmup=args.warmup, crop=list(args.crop), member=json.dumps(TRAIN_MEMBER),
                generation_s=t_gen)

    def save(records):
        net.save(args.out, epochs=len(records), records=json.dumps(records), **meta)

    t1 = time.time()
    records = train(net, xs, ts, args.epochs, args.lr, tuple(args.crop), seed=args.first_seed, val=(vx, vt),
                    time_budget=args.time_budget, on_epoch=save, warmup=args.warmup)
    print(json.dumps(dict(out=args.out, pairs=len(xs), epochs=len(records), first=records[0], last=records[-1],
                          generation_s=round(t_gen, 1), training_s=round(time.time() - t1, 1))), flush=True)
    del xs, ts
    torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
