"""Parity of the device-resident reconstruction (llr_recon_*) with the CPU
reference path (the oracle's FP32 restatement of lib.rs:227-341) and the FP64
oracle, on identical measurements and shift sequences.

Bar: relative NRMSE of the plain FP32 GPU image vs the FP32 CPU reference path
<= 1e-4.
"""
import numpy as np
import pytest

import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-4


def run_gpu(ds, meas=None, hybrid_used=None, **kw):
    from llr_hip import recon

    cfg = ds["cfg"]
    r = recon.Recon(ds["vol"], ds["nq"], ds["shifts"], blk=cfg["blk"], llr_mode=kw.get("llr_mode", "svt"),
                    rank=kw.get("rank", 0), lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"],
                    subspace=ds["subspace"], partial_blocks=kw.get("partial_blocks", False),
                    hybrid=kw.get("hybrid", True), fused_projection=kw.get("fused_projection"),
                    precision=kw.get("precision", "fp32"), storage=kw.get("storage", "fp32"),
                    residual_stop=kw.get("residual_stop", 0.0))
    if kw.get("stages") is not None:
        r.enable_timing(True)
    r.set_measurements(ds["meas"] if meas is None else meas)
    hist = r.run()
    img = r.image()
    lam, kc = r.info()
    if hybrid_used is not None:
        hybrid_used.append(r.hybrid())
    if kw.get("stages") is not None:
        kw["stages"].update(r.stage_times())
    r.close()
    return img, hist, lam, kc


def run_cpu(ds, dtype, **kw):
    cfg = ds["cfg"]
    sub = ds["subspace"]
    return O.recon(ds["meas"].astype(dtype), ds["vol"], ds["nq"], cfg["blk"], ds["shifts"],
                   llr_mode={"svt": 1, "rank": 0}[kw.get("llr_mode", "svt")], rank=kw.get("rank", 0),
                   lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"],
                   subspace=None if sub is None else sub.astype(dtype), partial_blocks=kw.get("partial_blocks", False))


def check(ds, label, **kw):
    img, hist, lam, kc = run_gpu(ds, **kw)
    x32, h32, lam32, kc32 = run_cpu(ds, np.complex64, **kw)
    x64, h64, lam64, kc64 = run_cpu(ds, np.complex128, **kw)
    e_ref = H.nrmse(img, x32)
    e_gpu64 = H.nrmse(img, x64)
    e_ref64 = H.nrmse(x32, x64)
    print(f"[{label}] NRMSE gpu-vs-cpu32 {e_ref:.3e}  gpu-vs-fp64 {e_gpu64:.3e}  cpu32-vs-fp64 {e_ref64:.3e}  "
          f"lambda gpu {lam:.6g} cpu {lam32:.6g}  final residual gpu {hist[-1]:.4e} cpu {h32[-1]:.4e}", flush=True)
    assert np.array_equal(kc, kc32) and np.array_equal(kc, kc64)
    assert np.all(np.isfinite(img))
    assert abs(lam - lam64) <= 1e-5 * lam64 + 1e-30
    assert np.allclose(hist, h64, rtol=1e-3)
    assert e_ref <= TOL, e_ref
    assert e_gpu64 <= TOL
    return e_ref


def test_recon_small_plain_svt(llr):
    check(H.make(H.SMALL_PLAIN), "32^3 x7 svt")


def test_recon_small_subspace_wavelet(llr):
    check(H.make(H.SMALL_SUBSPACE), "32x32x16 x13 subspace7 wavelet")


def test_recon_small_rank_mode(llr):
    ds = H.make(dict(H.SMALL_PLAIN, n_iter=4))
    check(ds, "32^3 x7 rank5", llr_mode="rank", rank=5)


def test_hybrid_iteration_matches_full_3d(llr):
    """Phase-encode masks: the (x, ky, kz) iteration (x passes skipped, measurements
    taken through the inverse x pass once) against the full 3D model and adjoint
    in every iteration, on the same data: equal up to FP32 rounding."""
    ds = H.make(H.SMALL_SUBSPACE)
    used = []
    img_h, hist_h, lam_h, _ = run_gpu(ds, hybrid_used=used)
    img_f, hist_f, lam_f, _ = run_gpu(ds, hybrid=False, hybrid_used=used)
    e = H.nrmse(img_h, img_f)
    print(f"[hybrid vs full 3D] NRMSE {e:.3e}; lambda {lam_h:.6g} / {lam_f:.6g}; residual {hist_h[-1]:.6e} / "
          f"{hist_f[-1]:.6e}", flush=True)
    assert used == [True, False]
    assert e <= 1e-5
    assert abs(lam_h - lam_f) <= 1e-5 * lam_f
    assert np.allclose(hist_h, hist_f, rtol=1e-4)


def test_hybrid_needs_whole_lines(llr):
    """One measured sample zeroed inside a measured x-line: the hard projection then
    keeps the estimate there, which the x-line form cannot express, so the set-up
    falls back to the full 3D iteration (bit-identical to asking for it)."""
    ds = H.make(H.SMALL_PLAIN)
    meas = ds["meas"].copy()
    nx = ds["vol"][0]
    line = int(np.flatnonzero(np.abs(meas) > 0)[0]) // nx
    meas[line * nx + nx // 3] = 0
    used = []
    img_a, hist_a, _, _ = run_gpu(ds, meas=meas, hybrid_used=used)
    img_f, hist_f, _, _ = run_gpu(ds, meas=meas, hybrid=False, hybrid_used=used)
    assert used == [False, False]
    assert np.array_equal(img_a, img_f) and np.array_equal(hist_a, hist_f)


def test_recon_config1_parity(llr):
    """Config 1: 64^3, 1 b0 + 6 directions, R=4, 8^3 blocks, SVT 0.02 sigma_max, 20 iterations."""
    ds = H.make("c1")
    check(ds, "c1 64^3 x7 20 it")


def test_recon_residual_decreases(llr):
    img, hist, lam, kc = run_gpu(H.make(H.SMALL_SUBSPACE))
    assert hist[-1] < hist[0]


@pytest.mark.parametrize("vol,blk", [((96, 48, 24), (10, 10, 10)), ((48, 24, 32), (6, 6, 6)),
                                     ((48, 24, 6), (6, 6, 6)), ((24, 48, 12), (8, 8, 8))])
def test_recon_partial_edge_blocks_radix3(llr, vol, blk):
    """Partial edge blocks on volumes with radix-3 FFT lengths (the uneven batch sizes).
    The subspace wavelet's z synthesis works in chunks of 16 samples: z lengths 6
    and 12 are one partial chunk at level 0, 24 a whole and a partial one, 32 two
    whole ones."""
    ds = H.make(dict(H.SMALL_SUBSPACE, vol=vol, blk=blk, n_iter=4))
    check(ds, f"{vol} blk {blk} partial", partial_blocks=True)


def test_recon_partial_equals_whole_when_divisible(llr):
    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=3))
    a = run_gpu(ds)[0]
    b = run_gpu(ds, partial_blocks=True)[0]
    assert np.array_equal(a, b)


@pytest.mark.timeout(600)
def test_recon_five_level_fused_subspace_wavelet(llr):
    """96^3 x13 with a rank-7 subspace and db2 0.005: the wavelet runs
    floor(log2(96/3)) = 5 levels, level 0 fused with the back-projection from the
    subspace coefficients (the 256x256x128 workload's path), 2 iterations."""
    ds = H.make(dict(H.SMALL_SUBSPACE, vol=(96, 96, 96), n_iter=2))
    check(ds, "96^3 x13 subspace7 wavelet 5 levels")


def test_recon_complex_subspace_matches_real(llr):
    """A complex subspace S Q (Q a seeded random 7x7 unitary) takes the general
    four-FMA expansion of the subspace wavelet; the real S takes the two-FMA
    one. X S Q and the block SVT commute with Q, and (C Q)(S Q)^H = C S^H, so
    both give the same image: each against the CPU path, and the two against
    each other."""
    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=3))
    rng = np.random.default_rng(77)
    K = ds["subspace"].shape[1]
    q, _ = np.linalg.qr(rng.standard_normal((K, K)) + 1j * rng.standard_normal((K, K)))
    dq = dict(ds, subspace=np.asfortranarray((ds["subspace"].astype(np.complex128) @ q).astype(np.complex64)))
    assert np.abs(dq["subspace"].imag).max() > 0.1
    e_c = check(dq, "32x32x16 x13 complex subspace")
    img_r = run_gpu(ds)[0]
    img_c = run_gpu(dq)[0]
    e = H.nrmse(img_c, img_r)
    print(f"[complex vs real subspace] NRMSE {e:.3e} (complex vs cpu32 {e_c:.3e})", flush=True)
    assert e <= 1e-5


WIDE = dict(H.SMALL_PLAIN, n_b0=7, dirs="repulsion", n_dirs=60, bval=3000.0, n_iter=3)


@pytest.mark.parametrize("mode,wav", [("svt", 0.0), ("svt", 0.005), ("rank", 0.0)])
def test_recon_no_subspace_67_volumes(llr, mode, wav):
    """The reference's default call without a subspace (lib.rs:290-292): the LLR
    step on 512 x 67 blocks of all volumes, through the wide kernels (Gram,
    workgroup Jacobi, apply), 32^3 x 67, against the CPU path."""
    ds = H.make(dict(WIDE, wavelet_alpha=wav))
    assert ds["nq"] == 67 and ds["subspace"] is None
    kw = dict(llr_mode="rank", rank=10) if mode == "rank" else {}
    check(ds, f"32^3 x67 no subspace {mode} wavelet {wav}", **kw)


def test_recon_no_subspace_partial_blocks_wide(llr):
    """Wide blocks with partial edge blocks (rows outside the volume are zero in
    the Gram and left alone by the apply)."""
    ds = H.make(dict(WIDE, vol=(48, 24, 32), blk=(10, 10, 10), n_iter=2))
    check(ds, "48x24x32 x67 no subspace partial 10^3", partial_blocks=True)


@pytest.mark.parametrize("vol,blk", [((32, 32, 16), (8, 8, 8)), ((48, 96, 24), (8, 8, 8)), ((24, 48, 12), (6, 6, 6))])
def test_fused_adjoint_projection_matches_unfused(llr, vol, blk):
    """The adjoint's inverse y pass handing the next LLR step C = X S directly
    (X formed once when the run ends) against storing X and projecting it in
    every iteration: the same image up to FP32 rounding, each against the CPU
    path. y lengths 32, 96 (radix 12 x 8) and 48 (12 x 4)."""
    ds = H.make(dict(H.SMALL_SUBSPACE, vol=vol, blk=blk, n_iter=4))
    kw = dict(partial_blocks=True)
    e_f = check(ds, f"{vol} fused adjoint projection", **kw)
    e_u = check(ds, f"{vol} stored image + projection", fused_projection=False, **kw)
    st_f, st_u = {}, {}
    img_f, hist_f = run_gpu(ds, stages=st_f, **kw)[:2]
    img_u, hist_u = run_gpu(ds, fused_projection=False, stages=st_u, **kw)[:2]
    e = H.nrmse(img_f, img_u)
    print(f"[{vol}] fused vs unfused NRMSE {e:.3e} (vs cpu32 {e_f:.3e} / {e_u:.3e}); launches fused "
          f"{st_f['fft_y_adjoint_project'][1]} + {st_f['fft_y_adjoint'][1]} y adjoint, "
          f"{st_f['subspace_project'][1]} projections", flush=True)
    assert e <= 1e-5
    assert np.allclose(hist_f, hist_u, rtol=1e-4)
    n = ds["cfg"]["n_iter"]
    # every adjoint but the last hands over coefficients; only the first LLR step projects
    assert st_f["fft_y_adjoint_project"][1] == n - 1 and st_f["fft_y_adjoint"][1] == 1
    assert st_u["fft_y_adjoint_project"][1] == 0 and st_u["fft_y_adjoint"][1] == n
    assert st_f["subspace_project"][1] < st_u["subspace_project"][1]


@pytest.mark.parametrize("storage", ["fp16", "fp32"])
def test_fused_adjoint_projection_mixed_form(llr, storage):
    """The mixed form (BF16 MFMA Gram; half or complex64 storage) with the fused
    adjoint projection (asked for: half storage does not take it by default):
    its FP32 projection of the unrounded adjoint replaces the BF16 one, so it
    is at least as close to the FP64 path as the unfused run, and a residual
    stop after a fused pass forms the image then (the image a run of that many
    iterations ends with, bit for bit)."""
    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=6))
    kw = dict(precision="mixed", storage=storage, fused_projection=True)
    x64 = run_cpu(ds, np.complex128)[0]
    img_f, hist_f = run_gpu(ds, **kw)[:2]
    img_u, hist_u = run_gpu(ds, **dict(kw, fused_projection=False))[:2]
    e_f, e_u = H.nrmse(img_f, x64), H.nrmse(img_u, x64)
    # stop after the third iteration (a fused adjoint): the image is formed then
    stop = float(hist_f[2]) * (1 + 1e-6)
    st = {}
    img_s, hist_s = run_gpu(ds, residual_stop=stop, stages=st, **kw)[:2]
    ds3 = dict(ds, shifts=ds["shifts"][:3])
    img_3 = run_gpu(ds3, **kw)[0]
    e_s = H.nrmse(img_s, img_3)
    print(f"[mixed {storage}] vs fp64: fused {e_f:.3e} unfused {e_u:.3e}; stop after 3: {len(hist_s)} iterations, "
          f"vs a 3-iteration run {e_s:.3e}", flush=True)
    assert e_f <= max(1.05 * e_u, 1e-4)
    assert len(hist_s) == 3 and st["fft_y_adjoint_project"][1] == 3 and st["fft_y_adjoint"][1] == 1
    # the deferred inverse y pass is the pass the 3-iteration run's last adjoint makes, on the same data
    assert np.array_equal(img_s, img_3), e_s


def test_fused_adjoint_projection_default_by_storage(llr):
    """By default the fused adjoint projection runs with complex64 storage and
    not with half storage (where the separate y pass and BF16 projection are
    faster)."""
    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=3))
    for storage, fused in (("fp32", True), ("fp16", False)):
        st = {}
        run_gpu(ds, precision="mixed", storage=storage, stages=st)
        assert (st["fft_y_adjoint_project"][1] > 0) == fused, (storage, st["fft_y_adjoint_project"])
