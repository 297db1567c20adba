"""Seeded synthetic diffusion data sets for the LLR reconstruction.

The reference ships no data generator; its inputs are scanner .cfl files. These
generators build what its loader would deliver, reusing the semantics of the
reference's helper fragments:

  - sample_unit_sphere (dti_subspace.rs:65-76): theta = 2*pi*u1, phi = acos(2*u2 - 1);
  - eigenvector_matrix (dti_subspace.rs:13-45): frame [e1, e1 x t, e1 x (e1 x t)];
  - sample_range (dti_subspace.rs:78-81);
  - the log-linear tensor design (diffusion_models.rs:221-253) plus an intercept
    column for the rank-7 subspace;
  - phase-encode masks of shape (ny, nz, nq) as under_sample_kspace expects
    (data_import.rs:416-438).

Everything is numpy and deterministic in the seeds. Arrays are flat,
column-major (x fastest, volume slowest), complex64.
"""
import numpy as np

PRESETS = {
    # 64^3, 1 b0 + 6 directions at b=1000, R=4, 20 iterations, no wavelet, no subspace
    "c1": dict(vol=(64, 64, 64), n_b0=1, dirs="six", n_dirs=6, bval=1000.0, R=4.0, n_iter=20, blk=(8, 8, 8),
               lambda_factor=0.02, wavelet_alpha=0.0, srank=0, n_ell=20, seeds=(1, 2, 3, 4)),
    # 256x256x128, 7 b0 + 60 directions at b=3000, rank-7 subspace, R=6, 30 iterations, db2 0.005
    "c2": dict(vol=(256, 256, 128), n_b0=7, dirs="repulsion", n_dirs=60, bval=3000.0, R=6.0, n_iter=30,
               blk=(8, 8, 8), lambda_factor=0.02, wavelet_alpha=0.005, srank=7, n_ell=20, seeds=(11, 12, 13, 14)),
}

BATCH_SIZES = [(96, 96, 64), (128, 128, 96), (192, 192, 96), (256, 256, 128)]
BATCH_VOLUMES = [13, 31, 67]
BATCH_R = [4.0, 6.0, 8.0]


def batch_member(seed):
    """One member of the uneven batch workload: matrix size, volume count and
    undersampling drawn from the seed; rank-7 subspace, db2 0.005, 30 iterations."""
    rng = np.random.default_rng(seed)
    vol = BATCH_SIZES[rng.integers(len(BATCH_SIZES))]
    nq = BATCH_VOLUMES[rng.integers(len(BATCH_VOLUMES))]
    R = BATCH_R[rng.integers(len(BATCH_R))]
    n_b0 = 7 if nq == 67 else 1
    return dict(vol=vol, n_b0=n_b0, dirs="repulsion", n_dirs=nq - n_b0, bval=3000.0, R=R, n_iter=30, blk=(8, 8, 8),
                lambda_factor=0.02, wavelet_alpha=0.005, srank=7, n_ell=20,
                seeds=(seed, seed + 1000, seed + 2000, seed + 3000))


def sample_unit_sphere(u1, u2):
    """dti_subspace.rs:65-76."""
    theta = 2.0 * np.pi * u1
    phi = np.arccos(2.0 * u2 - 1.0)
    return np.array([np.sin(phi) * np.cos(theta), np.sin(phi) * np.sin(theta), np.cos(phi)])


def eigenvector_matrix(principal_dir):
    """dti_subspace.rs:13-45: orthonormal frame whose first column is the principal direction."""
    e1 = np.asarray(principal_dir, np.float64)
    e1 = e1 / np.linalg.norm(e1)
    t = np.array([1.0, 0.0, 0.0])
    if abs(e1[0] - 1.0) < np.finfo(np.float32).eps:
        t = np.array([0.0, 1.0, 0.0])
    e2 = np.cross(e1, t)
    e2 = e2 / np.linalg.norm(e2)
    e3 = np.cross(e1, e2)
    return np.stack([e1, e2, e3], axis=1)


def six_directions():
    d = np.array([[1, 1, 0], [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]], np.float64)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def repulsion_directions(n, seed, iters=400):
    """n antipodally symmetric directions spread by electrostatic repulsion."""
    rng = np.random.default_rng(seed)
    u = rng.random((n, 2))
    d = np.stack([sample_unit_sphere(a, b) for a, b in u])
    step = 0.05
    for _ in range(iters):
        g = np.zeros_like(d)
        for s in (1.0, -1.0):
            diff = d[:, None, :] - s * d[None, :, :]
            r2 = np.sum(diff * diff, axis=-1)
            np.fill_diagonal(r2, np.inf)
            g += np.sum(diff / (r2[..., None] ** 1.5), axis=1)
        g -= np.sum(g * d, axis=1, keepdims=True) * d
        d = d + step * g / max(1.0, np.abs(g).max())
        d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d


def b_table(n_b0, dirs, bval):
    """b-values and unit gradient vectors, b0 volumes first."""
    bvals = np.concatenate([np.zeros(n_b0), np.full(len(dirs), float(bval))])
    bvecs = np.concatenate([np.zeros((n_b0, 3)), dirs], axis=0)
    return bvals, bvecs


def loglinear_design(bvals, bvecs):
    """[1, -b gx^2, -b gy^2, -b gz^2, -2b gx gy, -2b gx gz, -2b gy gz] (diffusion_models.rs:221-253 + intercept)."""
    b = np.asarray(bvals, np.float64)
    g = np.asarray(bvecs, np.float64)
    return np.stack([np.ones_like(b), -b * g[:, 0] ** 2, -b * g[:, 1] ** 2, -b * g[:, 2] ** 2,
                     -2 * b * g[:, 0] * g[:, 1], -2 * b * g[:, 0] * g[:, 2], -2 * b * g[:, 1] * g[:, 2]], axis=1)


def loglinear_subspace(bvals, bvecs, rank=7):
    """Orthonormal nq x rank basis of the log-linear design's column space
    (FP64 SVD), stored as complex64 with zero imaginary part (the .cfl matrix
    lib.rs:210-219 reads)."""
    A = loglinear_design(bvals, bvecs)
    U, s, _ = np.linalg.svd(A, full_matrices=False)
    r = min(rank, int(np.sum(s > s[0] * 1e-12)))
    return np.asfortranarray(U[:, :r].astype(np.complex64))


def fa_md_eigenvalues(fa, md):
    """Prolate tensor eigenvalues (l1, l2, l2) with the given FA and MD."""
    a = fa * np.sqrt(3.0 / (9.0 - 6.0 * fa * fa))
    return np.array([md * (1 + 2 * a), md * (1 - a), md * (1 - a)])


def phantom_params(seed, n_ell=20):
    """Ellipsoid geometry and tissue parameters (label 0 = outside, 1 = background)."""
    rng = np.random.default_rng(seed)
    ells = []
    for _ in range(n_ell):
        c = rng.uniform(-0.55, 0.55, 3)
        ax = rng.uniform(0.06, 0.25, 3)
        R = eigenvector_matrix(sample_unit_sphere(*rng.random(2)))
        pd = sample_unit_sphere(*rng.random(2))
        fa = rng.uniform(0.2, 0.8)
        md = rng.uniform(0.6e-3, 0.9e-3)
        s0 = rng.uniform(0.5, 1.0)
        E = eigenvector_matrix(pd)
        Dt = E @ np.diag(fa_md_eigenvalues(fa, md)) @ E.T
        ells.append(dict(c=c, ax=ax, R=R, D=Dt, s0=s0, fa=fa, md=md))
    bg_s0 = rng.uniform(0.5, 1.0)
    phase = dict(c0=rng.uniform(-np.pi, np.pi), lin=rng.uniform(-1.0, 1.0, 3), quad=rng.uniform(-0.5, 0.5, 6))
    return dict(ells=ells, bg_s0=bg_s0, bg_D=np.eye(3) * 3e-3, phase=phase)


def phantom_labels(vol, params):
    """Per-voxel label (uint8), x fastest; later ellipsoids overwrite earlier ones."""
    nx, ny, nz = vol
    u = [(np.arange(n) - n / 2) / (n / 2) for n in (nx, ny, nz)]
    X, Y, Z = np.meshgrid(u[0], u[1], u[2], indexing="ij")  # (nx, ny, nz)
    lab = np.zeros((nx, ny, nz), np.uint8)
    head = (X / 0.85) ** 2 + (Y / 0.9) ** 2 + (Z / 0.8) ** 2 <= 1.0
    lab[head] = 1
    P = np.stack([X, Y, Z], axis=-1)
    for k, e in enumerate(params["ells"]):
        q = (P - e["c"]) @ e["R"]
        inside = np.sum((q / e["ax"]) ** 2, axis=-1) <= 1.0
        lab[inside & head] = k + 2
    return lab


def phantom_phase(vol, params):
    nx, ny, nz = vol
    u = [(np.arange(n) - n / 2) / (n / 2) for n in (nx, ny, nz)]
    X, Y, Z = np.meshgrid(u[0], u[1], u[2], indexing="ij")
    p = params["phase"]
    ph = p["c0"] + p["lin"][0] * X + p["lin"][1] * Y + p["lin"][2] * Z
    q = p["quad"]
    ph = ph + q[0] * X * X + q[1] * Y * Y + q[2] * Z * Z + q[3] * X * Y + q[4] * X * Z + q[5] * Y * Z
    return np.exp(1j * ph)


def phantom_image(vol, bvals, bvecs, seed, n_ell=20, return_truth=False):
    """DTI phantom S(r, q) = S0(r) exp(-b g^T D(r) g) exp(i phi(r)); flat complex64, col-major."""
    params = phantom_params(seed, n_ell)
    lab = phantom_labels(vol, params)
    phase = phantom_phase(vol, params).astype(np.complex64)
    nq = len(bvals)
    N = int(np.prod(vol))
    D_list = [np.zeros((3, 3)), params["bg_D"]] + [e["D"] for e in params["ells"]]
    s0_list = [0.0, params["bg_s0"]] + [e["s0"] for e in params["ells"]]
    out = np.empty(N * nq, np.complex64)
    labf = lab.ravel(order="F")
    phf = phase.ravel(order="F")
    for qi in range(nq):
        g = bvecs[qi]
        amp = np.array([s0 * np.exp(-bvals[qi] * g @ D @ g) for s0, D in zip(s0_list, D_list)], np.float32)
        out[qi * N:(qi + 1) * N] = amp[labf] * phf
    if return_truth:
        return out, dict(labels=lab, params=params)
    return out


def pe_mask(ny, nz, nq, R, seed, centre=8):
    """Variable-density random phase-encode mask (ny, nz, nq), complex64 0/1,
    independent per volume, fully sampled centre x centre block, sampling
    fraction 1/R (data_import.rs:187-198 shape)."""
    rng = np.random.default_rng(seed)
    ky = (np.arange(ny) - ny / 2) / (ny / 2)
    kz = (np.arange(nz) - nz / 2) / (nz / 2)
    KY, KZ = np.meshgrid(ky, kz, indexing="ij")
    r = np.sqrt(KY ** 2 + KZ ** 2) / np.sqrt(2.0)
    cy, cz = ny // 2, nz // 2
    cen = np.zeros((ny, nz), bool)
    cen[cy - centre // 2:cy + centre // 2, cz - centre // 2:cz + centre // 2] = True
    target = 1.0 / R
    lo, hi = 0.0, 50.0
    for _ in range(60):  # density (1 - r)^p scaled to the target mean
        pw = 0.5 * (lo + hi)
        dens = np.where(cen, 1.0, np.clip((1 - r) ** pw, 0, 1))
        if dens.mean() > target:
            lo = pw
        else:
            hi = pw
    dens = np.where(cen, 1.0, np.clip((1 - r) ** (0.5 * (lo + hi)), 0, 1))
    m = np.empty((ny, nz, nq), np.complex64, order="F")
    for q in range(nq):
        m[:, :, q] = (rng.random((ny, nz)) < dens) | cen
    return m


def shift_sequence(vol, n_iter, seed):
    """One block shift per iteration, uniform in [0, vol) per axis (lib.rs:260, 420-427), seeded."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, v, n_iter) for v in vol], axis=1).astype(np.uint64)


def add_noise(ksp, sigma, seed):
    """Complex Gaussian noise with E|n|^2 = sigma^2 (in place on flat complex64)."""
    rng = np.random.default_rng(seed)
    n = ksp.size
    chunk = 1 << 24
    s = np.float32(sigma / np.sqrt(2.0))
    for i in range(0, n, chunk):
        j = min(n, i + chunk)
        nr = rng.standard_normal(j - i, np.float32)
        ni = rng.standard_normal(j - i, np.float32)
        ksp[i:j] += (nr + 1j * ni).astype(np.complex64) * s
    return ksp


def apply_mask(ksp, vol, nq, mask):
    """under_sample_kspace (data_import.rs:416-438) in numpy."""
    nx = vol[0]
    lines = ksp.reshape(-1, nx)
    m = np.asarray(mask).ravel(order="F")
    lines[m == 0] = 0
    return ksp


def dataset(preset, forward, vol=None, nq_override=None):
    """Build a measurement set. `forward(x, vol, nq)` must apply the forward
    signal model (signal_model.rs:34-79, zero corrections) in place on a flat
    complex64 array (the is_image path of lib.rs:151-168).

    Returns dict(meas, truth, subspace, shifts, mask, bvals, bvecs, cfg)."""
    cfg = dict(PRESETS[preset]) if isinstance(preset, str) else dict(preset)
    if vol is not None:
        cfg["vol"] = tuple(vol)
    s_ph, s_noise, s_mask, s_shift = cfg["seeds"]
    if cfg["dirs"] == "six":
        dirs = six_directions()
    else:
        dirs = repulsion_directions(cfg["n_dirs"], s_ph + 1000)
    bvals, bvecs = b_table(cfg["n_b0"], dirs, cfg["bval"])
    if nq_override:
        bvals, bvecs = bvals[:nq_override], bvecs[:nq_override]
    nq = len(bvals)
    v = cfg["vol"]
    img, truth = phantom_image(v, bvals, bvecs, s_ph, cfg["n_ell"], return_truth=True)
    s0_obj = float(np.mean(np.abs(img[:int(np.prod(v))][truth["labels"].ravel(order="F") > 0])))
    ksp = img.copy()
    forward(ksp, v, nq)
    add_noise(ksp, s0_obj / 30.0, s_noise)
    mask = pe_mask(v[1], v[2], nq, cfg["R"], s_mask)
    apply_mask(ksp, v, nq, mask)
    sub = loglinear_subspace(bvals, bvecs, cfg["srank"]) if cfg["srank"] else None
    shifts = shift_sequence(v, cfg["n_iter"], s_shift)
    return dict(meas=ksp, truth=img, subspace=sub, shifts=shifts, mask=mask, bvals=bvals, bvecs=bvecs, cfg=cfg,
                nq=nq, vol=v)


def device_dataset(preset, torch, keep_truth=False, log=None):
    """The measurement set of `preset` built on GPU 0, for workloads too large
    for the numpy path: the numpy phantom image, the library's forward signal
    model (signal_model.rs:34-79, zero corrections) on the device, complex
    Gaussian noise from a seeded torch generator, the phase-encode mask.
    Returns dict(y (device complex64), vol, nq, cfg, subspace, shifts, truth
    (device, if keep_truth), bvals, bvecs, mask)."""
    import time

    from . import signal_model

    say = log or (lambda *a: None)
    cfg = dict(PRESETS[preset]) if isinstance(preset, str) else dict(preset)
    s_ph, s_noise, s_mask, s_shift = cfg["seeds"]
    dirs = six_directions() if cfg["dirs"] == "six" else repulsion_directions(cfg["n_dirs"], s_ph + 1000)
    bvals, bvecs = b_table(cfg["n_b0"], dirs, cfg["bval"])
    nq = len(bvals)
    vol = tuple(cfg["vol"])
    N = int(np.prod(vol))
    t = time.time()
    img, geo = phantom_image(vol, bvals, bvecs, s_ph, cfg["n_ell"], return_truth=True)
    s0_obj = float(np.mean(np.abs(img[:N][geo["labels"].ravel(order="F") > 0])))
    say(f"phantom {vol} x{nq}: {time.time() - t:.1f} s")
    dev = torch.device("cuda", 0)
    x = torch.from_numpy(img).to(dev)
    truth = x.clone() if keep_truth else None
    del img
    torch.cuda.synchronize()
    zero = np.zeros((nq, 3), np.float32)
    signal_model.signal_model_device(x.data_ptr(), x.numel(), vol, nq, zero, zero, signal_model.ModelDirection.Forward)
    g = torch.Generator(device=dev)
    g.manual_seed(s_noise)
    sigma = s0_obj / 30.0
    noise = torch.randn(2 * x.numel(), device=dev, dtype=torch.float32, generator=g) * (sigma / np.sqrt(2.0))
    x += torch.view_as_complex(noise.view(-1, 2))
    del noise
    mask = pe_mask(vol[1], vol[2], nq, cfg["R"], s_mask)  # (ny, nz, nq)
    m = torch.from_numpy(np.ascontiguousarray(np.real(mask).transpose(2, 1, 0))).to(dev)  # (nq, nz, ny)
    xv = x.view(nq, vol[2], vol[1], vol[0])
    xv *= (m != 0).to(torch.complex64)[..., None]
    torch.cuda.synchronize()
    sub = loglinear_subspace(bvals, bvecs, cfg["srank"]) if cfg["srank"] else None
    shifts = shift_sequence(vol, cfg["n_iter"], s_shift)
    say(f"measurements on device: {time.time() - t:.1f} s, undersampling R={cfg['R']}")
    return dict(y=x, vol=vol, nq=nq, cfg=cfg, subspace=sub, shifts=shifts, truth=truth, bvals=bvals, bvecs=bvecs,
                mask=mask)
