"""GPU parity of every host-pointer entry point of libllr_hip.so against the oracle.

Integer/byte work (gather/scatter, shifts, masks, transposes) must be
bit-exact; floating point is compared with the tolerance written in each test
(FP32 kernels vs the FP64 oracle, relative to the data's norm).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(77)


def crandn(n, dtype=np.complex64):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)


def rel(a, b):
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(np.asarray(a, np.complex128) - b) / max(np.linalg.norm(b), 1e-300))


# ------------------------------------------------------------------ block.rs
@pytest.mark.parametrize("dtype", [np.float32, np.complex64, np.complex128, np.dtype("V3")])
@pytest.mark.parametrize("vol,blk,shift,start,n", [((10, 12, 14), (1, 1, 4), (0, 0, 12), 0, None),
                                                    ((64, 64, 64), (8, 8, 8), (61, 3, 40), 100, 300),
                                                    ((9, 7, 5), (3, 2, 2), (8, 6, 4), 3, None),
                                                    ((256, 256, 128), (8, 8, 8), (255, 129, 127), 16000, 384),
                                                    # 1024-voxel rows: 16-byte elements outgrow both LDS
                                                    # slab shapes and take the grid-row kernel
                                                    ((1024, 4, 4), (8, 2, 2), (5, 1, 3), 10, 200)])
def test_lift_unlift_bit_exact(llr, dtype, vol, blk, shift, start, n):
    from llr_hip import block

    nvols = 3
    N = int(np.prod(vol))
    dt = np.dtype(dtype)
    vols = rng.integers(0, 256, N * nvols * dt.itemsize, dtype=np.uint8).view(dt)
    g = block.grid_dim(vol, blk)
    total = int(np.prod(g))
    if n is None:
        n = total - start
    m = int(np.prod(blk))
    mats = np.zeros(n * nvols * m, vols.dtype)
    ref = np.zeros_like(mats)
    block.lift_data_set(vols, nvols, mats, vol, blk, shift, start, n)
    O.lift_data_set(vols, nvols, ref, vol, blk, shift, start, n)
    assert mats.tobytes() == ref.tobytes()
    # scatter into a scrambled copy: covered voxels restored, others untouched
    out = vols.copy()
    mats2 = np.zeros_like(mats)
    block.lift_data_set(out, nvols, mats2, vol, blk, shift, start, n)
    scram = np.ascontiguousarray(vols[::-1])
    refv = scram.copy()
    block.un_lift_data_set(scram, nvols, mats2, vol, blk, shift, start, n)
    O.un_lift_data_set(refv, nvols, mats2, vol, blk, shift, start, n)
    assert scram.tobytes() == refv.tobytes()


def test_lift_errors(llr):
    from llr_hip import block

    with pytest.raises(llr.LlrError):
        block.lift_data_set(np.zeros(8, np.float32), 1, np.zeros(64, np.float32), (2, 2, 2), (4, 1, 1), (0, 0, 0), 0, 1)
    with pytest.raises(llr.LlrError):  # matrix buffer too small (block.rs:152)
        block.lift_data_set(np.zeros(64, np.float32), 1, np.zeros(7, np.float32), (4, 4, 4), (2, 2, 2), (0, 0, 0), 0, 1)
    with pytest.raises(llr.LlrError):  # matrices past the block grid (8 blocks)
        block.lift_data_set(np.zeros(64, np.float32), 1, np.zeros(64, np.float32), (4, 4, 4), (2, 2, 2), (0, 0, 0), 7, 2)


@pytest.mark.parametrize("vol,blk,shift,start,n", [((64, 64, 64), (8, 8, 8), (61, 3, 40), 100, 300),
                                                    ((9, 7, 5), (3, 2, 2), (8, 6, 4), 3, 7)])
def test_lift_device_matches_host(llr, vol, blk, shift, start, n):
    """The device-pointer entry points give the host entry points' bytes."""
    import torch

    from llr_hip import block

    nvols = 5
    N = int(np.prod(vol))
    m = int(np.prod(blk))
    vols = (rng.standard_normal(N * nvols) + 1j * rng.standard_normal(N * nvols)).astype(np.complex64)
    ref = np.zeros(n * nvols * m, np.complex64)
    block.lift_data_set(vols, nvols, ref, vol, blk, shift, start, n)
    dev = torch.device("cuda", 0)
    dv = torch.from_numpy(vols).to(dev)
    dm = torch.zeros(n * nvols * m, dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()
    block.lift_data_set_device(dv.data_ptr(), nvols, 8, dm.data_ptr(), dm.numel(), vol, blk, shift, start, n)
    assert dm.cpu().numpy().tobytes() == ref.tobytes()
    out = np.ascontiguousarray(vols[::-1])
    refv = out.copy()
    block.un_lift_data_set(refv, nvols, ref, vol, blk, shift, start, n)
    do = torch.from_numpy(out).to(dev)
    torch.cuda.synchronize()
    block.un_lift_data_set_device(do.data_ptr(), nvols, 8, dm.data_ptr(), dm.numel(), vol, blk, shift, start, n)
    assert do.cpu().numpy().tobytes() == refv.tobytes()


# ----------------------------------------------------- cgemm / subspace / diag
def test_cgemm_kat(llr):
    from llr_hip import low_rank

    m, k, n, b = 4, 3, 4, 2
    A = np.zeros((b, k, m), np.complex64)
    B = np.zeros((b, n, k), np.complex64)
    for i in range(m):
        for j in range(k):
            A[:, j, i] = i + 1j * j
    for i in range(k):
        for j in range(n):
            B[:, j, i] = j + 1j * i
    C = np.zeros(b * n * m, np.complex64)
    low_rank.cublas_cgemm_strided_batched(m, k, n, b, False, 1, 0, False, False, A.ravel(), B.ravel(), C)
    expected = np.array([[-5, -5 + 3j, -5 + 6j, -5 + 9j], [-5 + 3j, -2 + 6j, 1 + 9j, 4 + 12j],
                         [-5 + 6j, 1 + 9j, 7 + 12j, 13 + 15j], [-5 + 9j, 4 + 12j, 13 + 15j, 22 + 18j]], np.complex64)
    C = C.reshape(b, n, m)
    assert np.array_equal(C[0].T, expected) and np.array_equal(C[0], C[1])


@pytest.mark.parametrize("ta,tb,bcast", [(True, True, True), (False, False, True), (False, True, False)])
def test_cgemm_vs_oracle(llr, ta, tb, bcast):
    from llr_hip import low_rank

    m, k, n, b = 7, 67, 33, 5
    A = crandn(m * k * (1 if bcast else b))
    B = crandn(k * n * b)
    C = crandn(m * n * b)
    Cr = C.astype(np.complex128)
    low_rank.cublas_cgemm_strided_batched(m, k, n, b, bcast, 1 + 0.5j, 0.25, ta, tb, A, B, C)
    O.cgemm(m, k, n, b, bcast, 1 + 0.5j, 0.25 + 0j, ta, tb, A.astype(np.complex128), B.astype(np.complex128), Cr)
    assert rel(C, Cr) < 2e-6


def test_diag_kat(llr):
    # cublas.rs:139-183 at reduced size: A = 1+1i, d_j = j+1 -> column j = (j+1)(1+i)
    from llr_hip import low_rank

    m, n, b = 2700, 67, 5
    A = np.full(m * n * b, 1 + 1j, np.complex64)
    d = np.tile(np.arange(1, n + 1, dtype=np.float32), b)
    low_rank.rhs_cdgmm_batched(m, n, b, A, d)
    expect = np.repeat(np.tile(np.arange(1, n + 1, dtype=np.float32), b), m) * (1 + 1j)
    assert np.array_equal(A, expect.astype(np.complex64))


def test_identity_subspace_kat(llr):
    from llr_hip import dti_subspace

    nq, rank, nvox, batch = 3, 3, 1000, 10
    S = np.eye(3, dtype=np.complex64).ravel(order="F")
    X = np.zeros((batch, nq, nvox), np.complex64)
    for j in range(nq):
        X[:, j, :] = np.arange(nvox) + j
    flat = X.ravel().copy()
    dti_subspace.subspace_constraint(S, flat, batch, nq, nvox, rank)
    assert np.array_equal(flat, X.ravel())


def test_subspace_constraint_vs_oracle(llr):
    """dti_subspace.rs:89-134 with a complex, non-orthonormal subspace and complex
    matrices: X' = conj(X) conj(S) S^T (the reference's final plain transpose),
    against the oracle, whose form is pinned by that closed form."""
    from llr_hip import dti_subspace

    for nq, srank, nvox, b in ((67, 7, 512, 5), (13, 3, 1000, 4)):
        S = crandn(nq * srank)
        X = crandn(b * nq * nvox)
        g = X.copy()
        dti_subspace.subspace_constraint(S, g, b, nq, nvox, srank)
        r = X.astype(np.complex128)
        O.subspace_constraint(S.astype(np.complex128), r, b, nq, nvox, srank)
        e = rel(g, r)
        print(f"[subspace_constraint nq {nq} srank {srank} nvox {nvox} x{b}] rel {e:.2e}", flush=True)
        assert e < 3e-6


def test_project_backproject_transpose_vs_oracle(llr):
    from llr_hip import dti_subspace

    nq, srank, nvox, b = 67, 7, 512, 9
    S = np.linalg.qr(crandn(nq * srank, np.complex128).reshape(nq, srank))[0].ravel(order="F").astype(np.complex64)
    X = crandn(b * nq * nvox)
    comp = np.zeros(b * srank * nvox, np.complex64)
    compr = np.zeros(b * srank * nvox, np.complex128)
    dti_subspace.dti_project(S, X, comp, b, nq, nvox, srank)
    O.dti_project(S.astype(np.complex128), X.astype(np.complex128), compr, b, nq, nvox, srank)
    assert rel(comp, compr) < 2e-6
    Xo = np.zeros_like(X)
    Xr = np.zeros(X.size, np.complex128)
    dti_subspace.dti_back_project(S, Xo, comp, b, nq, nvox, srank)
    O.dti_back_project(S.astype(np.complex128), Xr, comp.astype(np.complex128), b, nq, nvox, srank)
    assert rel(Xo, Xr) < 2e-6
    t = X.copy()
    dti_subspace.conj_transpose_matrix(t, nvox, nq, b)
    tr = X.copy()
    O.conj_transpose_matrix(tr, nvox, nq, b)
    assert np.array_equal(t, tr)


# ------------------------------------------------------------ low rank / SVT
def test_rank_one_ones(llr):
    from llr_hip import low_rank

    a = np.ones(20 * 7, np.complex64)
    low_rank.cu_low_rank_approx_batch(20, 7, 1, 1, a)
    assert np.abs(a - 1).max() < 1e-5


# n > 8 takes the wide kernels (Gram, workgroup Jacobi, apply); n > 96 keeps V in the work buffer
@pytest.mark.parametrize("m,n,r", [(512, 7, 3), (512, 8, 5), (343, 5, 5), (4096, 7, 2), (300, 12, 4), (512, 67, 7),
                                   (343, 16, 16), (300, 100, 10), (200, 128, 5)])
def test_low_rank_vs_oracle(llr, m, n, r):
    from llr_hip import low_rank

    b = 37
    # decaying spectrum so the truncation is well separated
    X = crandn(b * n * m).reshape(b, n, m) * (0.6 ** np.arange(n))[None, :, None]
    X = X.astype(np.complex64).ravel()
    g = X.copy()
    low_rank.cu_low_rank_approx_batch(m, n, r, b, g)
    ref = X.astype(np.complex128)
    O.low_rank_approx_batch(m, n, r, b, ref)
    assert rel(g, ref) < 2e-5


@pytest.mark.parametrize("m,n", [(512, 7), (512, 1), (729, 8), (200, 16), (512, 67), (1000, 97), (131, 9)])
def test_svt_vs_oracle(llr, m, n):
    from llr_hip import low_rank

    b = 41
    X = crandn(b * n * m)
    lam = 0.5 * np.sqrt(m)
    g = X.copy()
    smax = low_rank.svt_batch(m, n, lam, b, g)
    ref = X.astype(np.complex128)
    smax_ref = O.svt_batch(m, n, lam, b, ref)
    assert rel(g, ref) < 2e-5
    assert abs(smax - smax_ref) < 1e-5 * smax_ref


# ------------------------------------------------------------- signal model
@pytest.mark.parametrize("dims", [(64, 64, 64), (256, 256, 128), (8, 16, 4), (2, 512, 32), (128, 4, 256),
                                  (96, 48, 24), (192, 12, 6), (384, 3, 96), (6, 192, 12)])
def test_fftn_vs_numpy(llr, dims):
    from llr_hip import signal_model as sm

    nv = 2 if np.prod(dims) < 4_000_000 else 1
    x = crandn(nv * int(np.prod(dims)))
    y = x.copy()
    sm.cu_fftn_batch(dims, nv, sm.ModelDirection.Forward, sm.NormalizationType.NONE, y)
    ref = np.fft.fftn(x.astype(np.complex128).reshape(nv, *dims[::-1]), axes=(1, 2, 3)).ravel()
    assert rel(y, ref) < 1e-6
    sm.cu_fftn_batch(dims, nv, sm.ModelDirection.Inverse, sm.NormalizationType.Inverse, y)
    assert rel(y, x) < 1e-6


def test_fft_round_trip_ones_kat(llr):
    # cufft.rs:415-422 (cu_fft3 forward then inverse with 1/N on ones(128^3))
    from llr_hip import signal_model as sm

    x = np.ones(128 ** 3, np.complex64)
    sm.cu_fftn_batch((128, 128, 128), 1, 0, sm.NormalizationType.Inverse, x)
    sm.cu_fftn_batch((128, 128, 128), 1, 1, sm.NormalizationType.Inverse, x)
    assert np.abs(x - 1).max() < 1e-6


@pytest.mark.parametrize("dims", [(16, 8, 32), (64, 64, 64), (128, 32, 16), (96, 96, 48), (24, 192, 12)])
@pytest.mark.parametrize("direction", [0, 1])
def test_signal_model_vs_oracle(llr, dims, direction):
    from llr_hip import signal_model as sm

    nv = 3
    kc = np.array([[1, 0, -2], [0, 0, 0], [-1, 3, 1]], np.float32)
    ic = np.array([[0, 1, 0], [2, 0, 0], [0, 0, 0]], np.float32)
    x = crandn(nv * int(np.prod(dims)))
    g = x.copy()
    sm.signal_model(g, dims, nv, kc, ic, direction)
    ref = x.astype(np.complex128)
    O.signal_model(ref, dims, nv, kc, ic, direction)
    assert rel(g, ref) < 2e-6
    # against the reference's f32 ramps (the CPU reference path)
    ref32 = x.copy()
    O.signal_model(ref32, dims, nv, kc, ic, direction, phase_mode=0)
    assert rel(g, ref32) < 3e-5


def test_estimate_linear_correction(llr):
    from llr_hip import signal_model as sm

    dims = (64, 32, 16)
    k = crandn(int(np.prod(dims)))
    k[5 + 64 * (30 + 32 * 7)] = 50
    assert list(sm.estimate_linear_correction(k, dims)) == list(O.estimate_linear_correction(k, dims))
    k2 = np.zeros(int(np.prod(dims)), np.complex64)
    k2[[1000, 77, 5000]] = 3.0
    assert list(sm.estimate_linear_correction(k2, dims)) == list(O.estimate_linear_correction(k2, dims))


# -------------------------------------------------------------- fftshift.rs
@pytest.mark.parametrize("dims,dtype", [((5, 4, 7), np.complex64), ((256, 256, 128), np.complex64),
                                        ((3, 6, 1), np.float32), ((4, 4, 4), np.complex128), ((7, 3, 2), "V5")])
def test_fftshift_bit_exact(llr, dims, dtype):
    from llr_hip import fftshift

    n = int(np.prod(dims))
    a = rng.integers(0, 255, n * np.dtype(dtype).itemsize, dtype=np.uint8).view(dtype)
    b = a.copy()
    fftshift.fftshift(dims, b)
    r = a.copy()
    O.fftshift(r, dims)
    assert b.tobytes() == r.tobytes()
    fftshift.ifftshift(dims, b)
    assert b.tobytes() == a.tobytes()


def test_phase_shift3(llr):
    from llr_hip import fftshift

    dims = (32, 16, 8)
    x = crandn(int(np.prod(dims)))
    for d, sign in ((fftshift.PhaseShiftDir.Forward, -1), (fftshift.PhaseShiftDir.Inverse, 1)):
        g = x.copy()
        fftshift.phase_shift3(dims, g, d)
        r = x.copy()
        O.phase_shift3(r, dims, sign)
        # f32 phases up to ~88 rad: device sincosf vs host cos/sin differ by a few ulp of the angle
        assert rel(g, r) < 5e-6


# ------------------------------------------------------------------ wavelet
@pytest.mark.parametrize("dims", [(16, 12, 10), (15, 13, 11), (64, 64, 64), (256, 64, 32), (256, 256, 128)])
@pytest.mark.parametrize("alpha", [0.0, 0.3])
def test_wavelet_vs_oracle(llr, dims, alpha):
    from llr_hip import recon

    nv = 2
    x = crandn(nv * int(np.prod(dims)))
    g = x.copy()
    recon.wavelet_denoise(g, nv, dims, alpha)
    ref = x.astype(np.complex128)
    N = int(np.prod(dims))
    for v in range(nv):
        seg = ref[v * N:(v + 1) * N].copy()
        O.wavelet_denoise(seg, dims, alpha)
        ref[v * N:(v + 1) * N] = seg
    assert rel(g, ref) < 3e-6
    if alpha == 0.0:
        assert rel(g, x) < 3e-6


# ---------------------------------------------------- data consistency helpers
def test_hard_project_and_stats(llr):
    from llr_hip import recon

    n = 3_000_017
    x = crandn(n)
    y = crandn(n)
    y[rng.random(n) < 0.7] = 0
    xg = x.copy()
    e = recon.hard_project(xg, y)
    xr = x.copy()
    er = O.hard_project(xr, y)
    assert np.array_equal(xg, xr)
    assert abs(e - er) < 1e-9 * er
    us, en = recon.calc_undersampling_factor(y)
    usr, enr = O.calc_undersampling_factor(y)
    assert us == usr and abs(en - enr) < 1e-9 * enr


def test_under_sample_kspace(llr):
    from llr_hip import recon

    dims, nq = (32, 16, 8), 3
    d = crandn(int(np.prod(dims)) * nq)
    mask = (rng.random(16 * 8 * nq) < 0.4).astype(np.complex64)
    g = d.copy()
    recon.under_sample_kspace(g, dims, nq, mask)
    r = d.copy()
    O.under_sample_kspace(r, dims, nq, mask)
    assert np.array_equal(g, r)


# --------------------------------------------------------------- DTI scalars
@pytest.mark.parametrize("noise", [0.0, 0.05])
def test_dti_scalars_vs_oracle(llr, noise):
    from llr_hip import diffusion_models as dm
    from llr_hip import phantom

    vol = (24, 20, 16)
    N = int(np.prod(vol))
    bvals, bvecs = phantom.b_table(2, phantom.repulsion_directions(15, 3), 1000.0)
    img = phantom.phantom_image(vol, bvals, bvecs, 5, 8)
    if noise:
        img = (img + noise * crandn(img.size)).astype(np.complex64)
    got = dm.calc_scalars(img, N, bvals, bvecs)
    ref = O.dti_scalars(img, N, bvals, bvecs)
    for i, k in enumerate(dm.SCALAR_NAMES):
        g, r = got[k].astype(np.float64), ref[i]
        assert np.array_equal(np.isnan(g), np.isnan(r)), k
        ok = ~np.isnan(r)
        scale = np.abs(ref[:3]).max() if k != "fa" else 1.0
        assert np.abs(g[ok] - r[ok]).max() <= 2e-6 * scale, k
