"""Mixed-precision form (BF16 MFMA Gram and projection with FP32 accumulation,
one FP32 Jacobi refinement sweep, FP64 eigen-solve for blocks whose Gram
condition number exceeds 1e6, FP64 residual) against the FP64 oracle.

Bars (relative NRMSE vs the FP64 oracle on the same measurements and shifts):
without a subspace the BF16 products only seed the eigenvectors, which the
FP32 sweep refines, so the image stays at FP32 accuracy (<= 1e-4); with a
subspace the BF16 projection and back-projection round every coefficient to
8 bits of mantissa, bounded here by 5e-3. Config 3 also reports the per-voxel
FA / MD error of the tensor fit (diffusion_models.rs:21-129) inside the object.

Half storage (iterate, measurements' working copy and wavelet workspace held
as IEEE FP16 pairs at the scale sqrt(voxels / energy), every kernel computing
in FP32): each pass rounds what it stores to 11 significant bits (relative
2^-11), about ten roundings per iteration, none accumulating past the next
data-consistency projection. The mixed form's BF16 back-projection dominates
its error either way; bounded by 5e-3 relative NRMSE vs the FP64 oracle and
2e-2 on the residuals.
"""
import numpy as np
import pytest

import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu


def run(ds, precision, storage="fp32", **kw):
    from llr_hip import recon

    cfg = ds["cfg"]
    r = recon.Recon(ds["vol"], ds["nq"], ds["shifts"], blk=cfg["blk"], llr_mode=kw.get("llr_mode", "svt"),
                    rank=kw.get("rank", 0), lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"],
                    subspace=ds["subspace"], precision=precision, storage=storage)
    r.set_measurements(ds["meas"])
    hist = r.run()
    img = r.image()
    lam, _ = r.info()
    stats = r.mixed_stats()
    r.close()
    return img, hist, lam, stats


def fp64(ds, **kw):
    cfg = ds["cfg"]
    sub = ds["subspace"]
    return O.recon(ds["meas"].astype(np.complex128), ds["vol"], ds["nq"], cfg["blk"], ds["shifts"],
                   llr_mode={"svt": 1, "rank": 0}[kw.get("llr_mode", "svt")], rank=kw.get("rank", 0),
                   lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"],
                   subspace=None if sub is None else sub.astype(np.complex128))


def test_mixed_plain_small_matches_fp64(llr):
    ds = H.make(H.SMALL_PLAIN)
    img, hist, lam, (nf, nb) = run(ds, "mixed")
    x64, h64, lam64, _ = fp64(ds)
    e = H.nrmse(img, x64)
    print(f"[mixed 32^3 x7] NRMSE vs fp64 {e:.3e}, fp64 eigen-solves {nf}/{nb}", flush=True)
    assert nb == len(ds["shifts"]) * 64
    assert abs(lam - lam64) <= 1e-4 * lam64
    assert np.allclose(hist, h64, rtol=1e-3)
    assert e <= 1e-4


def test_mixed_subspace_wavelet_matches_fp64(llr):
    ds = H.make(H.SMALL_SUBSPACE)
    img, hist, lam, (nf, nb) = run(ds, "mixed")
    x64, h64, lam64, _ = fp64(ds)
    plain = run(ds, "fp32")[0]
    e, ep = H.nrmse(img, x64), H.nrmse(plain, x64)
    print(f"[mixed 32x32x16 x13 subspace] NRMSE vs fp64 {e:.3e} (plain fp32 {ep:.3e}), fp64 eigen-solves {nf}/{nb}",
          flush=True)
    assert np.all(np.isfinite(img))
    assert abs(lam - lam64) <= 1e-2 * lam64
    assert np.allclose(hist, h64, rtol=2e-2)
    assert e <= 5e-3


def test_mixed_rank_mode(llr):
    ds = H.make(dict(H.SMALL_PLAIN, n_iter=4))
    img = run(ds, "mixed", llr_mode="rank", rank=5)[0]
    x64 = fp64(ds, llr_mode="rank", rank=5)[0]
    assert H.nrmse(img, x64) <= 1e-4


def test_config3_mixed_fa_md_error(llr):
    """Config 3: config-1 inputs through the mixed form; image NRMSE and per-voxel FA/MD
    error inside the object against the FP64 oracle reconstruction."""
    from llr_hip import diffusion_models as dm
    from llr_hip import phantom

    ds = H.make("c1")
    img, hist, lam, (nf, nb) = run(ds, "mixed")
    x64 = fp64(ds)[0]
    e = H.nrmse(img, x64)
    N = int(np.prod(ds["vol"]))
    got = dm.calc_scalars(img, N, ds["bvals"], ds["bvecs"])
    ref = O.dti_scalars(x64, N, ds["bvals"], ds["bvecs"])
    params = phantom.phantom_params(ds["cfg"]["seeds"][0], ds["cfg"]["n_ell"])
    obj = phantom.phantom_labels(ds["vol"], params).ravel(order="F") > 0
    ok = obj & np.isfinite(ref[3]) & np.isfinite(got["fa"])
    dfa = np.abs(got["fa"][ok] - ref[3][ok])
    dmd = np.abs(got["md"][ok] - ref[4][ok]) / np.maximum(np.abs(ref[4][ok]), 1e-12)
    print(f"[config 3] image NRMSE vs fp64 {e:.3e}; FA |err| mean {dfa.mean():.2e} max {dfa.max():.2e}; "
          f"MD rel err mean {dmd.mean():.2e} p99 {np.percentile(dmd, 99):.2e}; fp64 eigen-solves {nf}/{nb}", flush=True)
    assert e <= 1e-4
    assert dfa.mean() <= 1e-4 and np.percentile(dmd, 99) <= 1e-3


def test_mixed_fp64_fallback_on_rank_deficient_blocks(llr):
    """Volumes that are scaled copies of one image make every block's Gram matrix
    rank one (condition number ~1e14): the FP64 eigen-solve must take over."""
    from llr_hip import phantom

    vol, nq = (32, 32, 16), 7
    N = int(np.prod(vol))
    base = phantom.phantom_image(vol, np.zeros(1), np.zeros((1, 3)), 41, 6)
    scale = np.linspace(1.0, 0.4, nq) * np.exp(1j * np.linspace(0, 2, nq))
    img = (scale[:, None] * base[None, :]).astype(np.complex64).ravel()
    meas = img.copy()
    H.oracle_forward(meas, vol, nq)
    shifts = phantom.shift_sequence(vol, 3, 43)
    ds = dict(meas=meas, vol=vol, nq=nq, shifts=shifts, subspace=None,
              cfg=dict(blk=(8, 8, 8), lambda_factor=0.02, wavelet_alpha=0.0))
    out, hist, lam, (nf, nb) = run(ds, "mixed")
    x64 = fp64(ds)[0]
    e = H.nrmse(out, x64)
    print(f"[mixed rank-one blocks] NRMSE vs fp64 {e:.3e}, fp64 eigen-solves {nf}/{nb}", flush=True)
    assert nf > nb // 4
    assert e <= 1e-5


@pytest.mark.parametrize("vol", [(32, 32, 16), (48, 24, 16), (96, 96, 96)])
def test_fp16_storage_matches_fp64(llr, vol):
    """Mixed form with half storage against the FP64 oracle, radix-2 and radix-3 sizes."""
    ds = H.make(dict(H.SMALL_SUBSPACE, vol=vol))
    print(f"[fp16 storage {vol} x{ds['nq']}] running", flush=True)
    img, hist, lam, _ = run(ds, "mixed", storage="fp16")
    x64, h64, lam64, _ = fp64(ds)
    img32 = run(ds, "mixed")[0]
    e, e32 = H.nrmse(img, x64), H.nrmse(img32, x64)
    print(f"[fp16 storage {vol} x{ds['nq']}] NRMSE vs fp64 {e:.3e} (fp32 storage {e32:.3e}); "
          f"residual {hist[-1]:.5e} vs {h64[-1]:.5e}", flush=True)
    assert np.all(np.isfinite(img))
    assert abs(lam - lam64) <= 2e-2 * lam64
    assert np.allclose(hist, h64, rtol=2e-2)
    assert e <= 5e-3


@pytest.mark.parametrize("gain", [1.0, 3.0e4, 1.0e-4])
def test_fp16_storage_round_trips(llr, gain):
    """Initial image in, image out at any data scale: half-precision rounding
    only (relative 2^-11 per part, the stored data being scaled to O(1));
    coefficients of the stored iterate match an FP64 projection of the rounded
    image (the hi + lo BF16 MFMA projection carries X and S to ~2^-16) and come
    back in the caller's units; set_coefficients scales them back in."""
    from llr_hip import recon

    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=1))
    cfg = ds["cfg"]
    vol, nq, K = ds["vol"], ds["nq"], cfg["srank"]
    N = int(np.prod(vol))
    rng = np.random.default_rng(5)
    x = (gain * (rng.standard_normal(N * nq) + 1j * rng.standard_normal(N * nq))).astype(np.complex64)
    r = recon.Recon(vol, nq, ds["shifts"], blk=cfg["blk"], lambda_factor=cfg["lambda_factor"],
                    wavelet_alpha=cfg["wavelet_alpha"], subspace=ds["subspace"], precision="mixed", storage="fp16")
    r.set_measurements((gain * ds["meas"]).astype(np.complex64))  # sets the storage scale
    r.set_initial(x)
    r.prepare()
    out = r.image()
    # stored values are x * hs, hs = sqrt(voxels / measurement energy); parts that
    # land among half's subnormals (below 2^-14 once scaled) keep an absolute
    # precision of 2^-24 / hs instead of a relative one
    meas = (gain * ds["meas"]).astype(np.complex64)
    hs = np.sqrt(N / np.sum(np.abs(meas.astype(np.complex128)) ** 2))
    xr, xi = np.abs(x.real) * hs, np.abs(x.imag) * hs
    er, ei = np.abs(out.real - x.real) * hs, np.abs(out.imag - x.imag) * hs
    normal = np.concatenate([xr >= 2.0 ** -14, xi >= 2.0 ** -14])
    err = np.concatenate([er, ei])
    mag = np.concatenate([xr, xi])
    rel = err[normal] / mag[normal]
    assert np.all(np.isfinite(out))
    assert rel.max() <= 2.0 ** -11 * 1.01 and err[~normal].max(initial=0.0) <= 2.0 ** -24
    c = r.coefficients().reshape(K, N)
    S = np.asarray(ds["subspace"], np.complex128)  # nq x K
    ref = (S.T @ out.astype(np.complex128).reshape(nq, N))  # C[k][v] = sum_q X[q][v] S[q][k]
    e = H.nrmse(c, ref)
    # coefficients in, image out: X = C S^H, in the caller's units
    r.set_coefficients(c.ravel())
    back = r.image()
    r.close()
    e2 = H.nrmse(back, (np.conj(S) @ c.astype(np.complex128)).ravel())
    print(f"[fp16 storage round trip, gain {gain:g}] max rel rounding {rel.max():.3e}; projection NRMSE {e:.3e}; "
          f"back-projection NRMSE {e2:.3e}", flush=True)
    assert e <= 1e-4 and e2 <= 1e-3


def _fp16_recon(ds, storage="fp16", hybrid=True, n_iter=None):
    from llr_hip import recon

    cfg = ds["cfg"]
    sh = ds["shifts"] if n_iter is None else ds["shifts"][:n_iter]
    return recon.Recon(ds["vol"], ds["nq"], sh, blk=cfg["blk"], lambda_factor=cfg["lambda_factor"],
                       wavelet_alpha=cfg["wavelet_alpha"], subspace=ds["subspace"], precision="mixed",
                       storage=storage, hybrid=hybrid)


def test_fp16_set_coefficients_before_prepare(llr):
    """Coefficients set before any set-up (the storage scale is not known yet)
    wait in complex64 and are converted at the set-up's scale: the image after
    prepare() is C S^H to half rounding, in the caller's units."""
    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=1))
    vol, nq, K = ds["vol"], ds["nq"], ds["cfg"]["srank"]
    N = int(np.prod(vol))
    rng = np.random.default_rng(8)
    c = (30.0 * (rng.standard_normal(K * N) + 1j * rng.standard_normal(K * N))).astype(np.complex64)
    S = np.asarray(ds["subspace"], np.complex128)
    expect = (np.conj(S) @ c.astype(np.complex128).reshape(K, N)).ravel()
    for storage in ("fp16", "fp32"):
        r = _fp16_recon(ds, storage)
        r.set_measurements(ds["meas"])
        r.set_coefficients(c)
        r.prepare()
        img = r.image()
        r.close()
        e = H.nrmse(img, expect)
        print(f"[set_coefficients before prepare, {storage}] NRMSE vs C S^H {e:.3e}", flush=True)
        assert e <= (1e-3 if storage == "fp16" else 1e-6)


def test_fp16_set_initial_after_prepare(llr):
    """An initial image set after prepare() is the start of the next run (which
    skips the set-up) in half storage as in complex64 storage."""
    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=2))
    N = int(np.prod(ds["vol"])) * ds["nq"]
    rng = np.random.default_rng(9)
    x = (0.3 * (rng.standard_normal(N) + 1j * rng.standard_normal(N))).astype(np.complex64)
    out = {}
    for storage in ("fp16", "fp32"):
        r = _fp16_recon(ds, storage)
        r.set_measurements(ds["meas"])
        r.prepare()
        r.set_initial(x)
        start = r.image()
        r.run()
        out[storage] = r.image()
        r.close()
        e0 = H.nrmse(start, x)
        print(f"[set_initial after prepare, {storage}] start vs x {e0:.3e}", flush=True)
        assert e0 <= (1e-3 if storage == "fp16" else 0.0)
    cold = _fp16_recon(ds, "fp32")
    cold.set_measurements(ds["meas"])
    cold.run()
    img_cold = cold.image()
    cold.close()
    e = H.nrmse(out["fp16"], out["fp32"])
    ec = H.nrmse(out["fp32"], img_cold)
    print(f"[set_initial after prepare] fp16 vs fp32 storage {e:.3e}; warm vs cold start {ec:.3e}", flush=True)
    assert e <= 5e-3 and ec > 10 * e


def test_fp16_full_3d_matches_fp64(llr):
    """Half storage with the full 3D iteration: asked for (no_hybrid), and forced by
    measurements whose x-lines are partly measured (single samples zeroed)."""
    ds = H.make(H.SMALL_SUBSPACE)
    x64 = fp64(ds)[0]
    r = _fp16_recon(ds, hybrid=False)
    r.set_measurements(ds["meas"])
    h = r.run()
    img = r.image()
    r.close()
    e = H.nrmse(img, x64)
    print(f"[fp16 full 3D, no_hybrid] NRMSE vs fp64 {e:.3e}", flush=True)
    assert e <= 5e-3
    meas = ds["meas"].copy()
    rng = np.random.default_rng(10)
    meas[rng.random(meas.size) < 0.1] = 0
    ds2 = dict(ds, meas=meas)
    x64b = fp64(ds2)[0]
    r = _fp16_recon(ds2)
    r.set_measurements(meas)
    r.run()
    used = r.hybrid()
    img = r.image()
    r.close()
    e2 = H.nrmse(img, x64b)
    print(f"[fp16 full 3D, samples zeroed inside lines] hybrid {used}; NRMSE vs fp64 {e2:.3e}", flush=True)
    assert not used and e2 <= 5e-3


def test_fp16_keeps_tiny_measured_samples(llr):
    """A measured sample whose half-precision value would round to zero stays
    measured (y != 0 is the hard projection's sampled set, lib.rs:396): 5 % of
    the measured samples scaled by 1e-9, full 3D iteration, half against
    complex64 storage of the same mixed form."""
    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=3))
    meas = ds["meas"].copy()
    rng = np.random.default_rng(11)
    nz = np.flatnonzero(meas)
    tiny = nz[rng.random(nz.size) < 0.05]
    meas[tiny] *= np.float32(1e-9)
    imgs = {}
    for storage in ("fp16", "fp32"):
        r = _fp16_recon(dict(ds, meas=meas), storage, hybrid=False)
        r.set_measurements(meas)
        r.run()
        imgs[storage] = r.image()
        r.close()
    e = H.nrmse(imgs["fp16"], imgs["fp32"])
    print(f"[fp16 tiny measured samples, {tiny.size} of {nz.size}] NRMSE fp16 vs fp32 storage {e:.3e}", flush=True)
    assert e <= 5e-3
