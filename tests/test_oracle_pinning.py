"""Pins the oracle's reconstruction loop and multi-level wavelet with checks
that do not reuse the oracle's own code.

  - An independent numpy FP64 restatement of llr_recon_exec's loop
    (lib.rs:262-341), written from the reference sources: the signal model
    from signal_model.rs:34-117 and cufft.rs:320-326, the block gather and
    scatter from block.rs:8-184 by index formula, the low-rank step from a
    LAPACK SVD, and the wavelet from explicit analysis / synthesis matrices
    built from the single-level formulas of kernels.cu:155-268. Two
    iterations at tiny sizes must match the oracle to FP64 rounding, and the
    same restatement with LLR and wavelet swapped, or with the residual not
    divided by the total energy (lib.rs:321), must not.
  - Closed forms of the loop: with every sample measured the image after each
    iteration is A^H y exactly; with an identity proximal step (rank = all
    columns, no wavelet) the zero-filled adjoint is a fixed point and the
    residual is 0.
  - The multi-level wavelet against the explicit-matrix construction: level
    rule floor(log2(min dim / 3)), subband set (every detail band and the
    deepest approximation shrunk), odd-length bookkeeping. The un-vendored
    dwt crate behind lib.rs:439-454 is not available: its own conventions stay
    unpinned; these checks pin the oracle to the pywt-style reading of them.
"""
import numpy as np
import pytest

import oracle as O

rng = np.random.default_rng(2024)


def crandn(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# ------------------------------------------------------------ signal model
def to_zyx(x, vol, nq):
    return x.reshape(nq, vol[2], vol[1], vol[0])


def ramp(vol, corr, sign):
    """phase_shift3_sub_vox (cufft.rs:290-326) with s = c + n/2 (calc_shift,
    signal_model.rs:110-117), exact phase: exp(i sign sum_a (i_a - n_a/2) 2 pi s_a / n_a),
    as a (nz, ny, nx) array."""
    ph = 0.0
    for a, n in enumerate(vol):
        s = float(corr[a]) + 0.5 * n
        i = np.arange(n) - 0.5 * n
        shape = [1, 1, 1]
        shape[2 - a] = n
        ph = ph + (i * 2.0 * np.pi * s / n).reshape(shape)
    return np.exp(1j * sign * ph)


def signal_model(x, vol, nq, kc, ic, forward):
    """signal_model.rs:34-79: forward = ramp(ksp, +) -> unitary forward FFT -> ramp(img, -);
    inverse = ramp(img, +) -> unitary inverse FFT -> ramp(ksp, -)."""
    v = to_zyx(x.astype(np.complex128), vol, nq).copy()
    for q in range(nq):
        if forward:
            v[q] = np.fft.fftn(v[q] * ramp(vol, kc[q], +1), norm="ortho") * ramp(vol, ic[q], -1)
        else:
            v[q] = np.fft.ifftn(v[q] * ramp(vol, ic[q], +1), norm="ortho") * ramp(vol, kc[q], -1)
    return v.ravel()


def linear_corrections(y, vol, nq):
    """estimate_linear_correction (signal_model.rs:100-106): arg-max |k|^2 (lowest
    index on ties) minus (n + 1) / 2 per axis."""
    out = np.zeros((nq, 3))
    N = int(np.prod(vol))
    for q in range(nq):
        idx = int(np.argmax(np.abs(y[q * N:(q + 1) * N]) ** 2))
        sub = (idx % vol[0], (idx // vol[0]) % vol[1], idx // (vol[0] * vol[1]))
        out[q] = [sub[a] - (vol[a] + 1) // 2 for a in range(3)]
    return out


# ------------------------------------------------------------ blocks (block.rs)
def block_addresses(vol, blk, shift):
    """Voxel index of row i of block b: block corner sub2ind(b, grid) * blk + shift,
    wrapped (block.rs:8-28, 112-133), rows x fastest inside the block."""
    grid = [vol[a] // blk[a] for a in range(3)]
    nb = grid[0] * grid[1] * grid[2]
    m = blk[0] * blk[1] * blk[2]
    addr = np.empty((nb, m), np.int64)
    for b in range(nb):
        c = (b % grid[0], (b // grid[0]) % grid[1], b // (grid[0] * grid[1]))
        for i in range(m):
            d = (i % blk[0], (i // blk[0]) % blk[1], i // (blk[0] * blk[1]))
            p = [(c[a] * blk[a] + shift[a] + d[a]) % vol[a] for a in range(3)]
            addr[b, i] = p[0] + vol[0] * (p[1] + vol[1] * p[2])
    return addr


def lowrank(A, mode, lam, rank):
    """cu_low_rank_approx_batch (low_rank.rs:10-69): U_r diag(s_r) V_r^H, or the soft threshold."""
    U, s, Vh = np.linalg.svd(A, full_matrices=False)
    g = s.copy()
    if mode == "svt":
        g = np.maximum(s - lam, 0.0)
    else:
        g[rank:] = 0.0
    return (U * g) @ Vh, s[0]


def llr_step(x, vol, nq, blk, shift, S, mode, lam, rank, modify=True):
    """lib.rs:270-296: gather, [X S], low rank, [L S^H], scatter. Returns sigma_max seen."""
    N = int(np.prod(vol))
    addr = block_addresses(vol, blk, shift)
    X = x.reshape(nq, N)
    smax = 0.0
    for b in range(addr.shape[0]):
        M = X[:, addr[b]].T  # m x nq Casorati matrix (column q = volume q)
        T = M @ S if S is not None else M
        L, s0 = lowrank(T, mode, lam, rank)
        smax = max(smax, s0)
        if modify:
            X[:, addr[b]] = (L @ S.conj().T if S is not None else L).T
    return smax


# ------------------------------------------------------------ wavelet matrices
H = np.array([1 + np.sqrt(3), 3 + np.sqrt(3), 3 - np.sqrt(3), 1 - np.sqrt(3)]) / (4 * np.sqrt(2))
LO_D, HI_D = H[::-1].copy(), np.array([-H[0], H[1], -H[2], H[3]])  # pywt db2 dec_lo / dec_hi
LO_R, HI_R = H.copy(), HI_D[::-1].copy()                               # rec_lo / rec_hi


def sym(i, n):
    """symmetric_boundary_index (kernels.cu:97-105), applied until inside."""
    while i < 0 or i >= n:
        i = -i - 1 if i < 0 else 2 * n - i - 1
    return i


def analysis_matrix(n, F=4):
    """_dwt3_axis (kernels.cu:155-212) as a (2 nc x n) matrix: coefficient c reads
    sample sym(2c - (F-1) + j + 1) with filter tap F-1-j; approximation rows
    [0, nc), detail rows [nc, 2 nc); nc = (n + F - 1) / 2 (kernels.cu:150-152)."""
    nc = (n + F - 1) // 2
    W = np.zeros((2 * nc, n))
    for c in range(nc):
        for j in range(F):
            k = sym(2 * c - (F - 1) + j + 1, n)
            W[c, k] += LO_D[F - 1 - j]
            W[nc + c, k] += HI_D[F - 1 - j]
    return W


def synthesis_matrix(n, nc, F=4):
    """_idwt3_axis (kernels.cu:215-268) as an (n x 2 nc) matrix: sample s sums
    rec taps j over even idx = start + s - j inside [0, 2 nc), start =
    (full - keep) / 2 with full = 2 nc + F - 1, keep = 2 nc - F + 2."""
    full, keep = 2 * nc + F - 1, 2 * nc - F + 2
    start = (full - keep) // 2
    Sm = np.zeros((n, 2 * nc))
    for s in range(n):
        for j in range(F):
            idx = start + s - j
            if 0 <= idx < 2 * nc and idx % 2 == 0:
                Sm[s, idx // 2] += LO_R[j]
                Sm[s, idx // 2 + nc] += HI_R[j]
    return Sm


def apply3(box, mats):
    """Apply one matrix per axis to a (nz, ny, nx) box (mats in x, y, z order)."""
    return np.einsum("zi,yj,xk,ijk->zyx", mats[2], mats[1], mats[0], box, optimize=True)


def shrink(v, alpha):
    """shrink_complex (lib.rs:460-465), 0 for |x| = 0."""
    m = np.abs(v)
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.where(m > 0, np.maximum(m - alpha, 0.0) / np.where(m > 0, m, 1.0), 0.0)
    return v * f


def wavelet_levels(vol):
    """floor(log2(min dim / (F - 1))) with F = 4 (the pywt dwt_max_level rule)."""
    n = min(vol)
    return int(np.floor(np.log2(n / 3.0))) if n >= 3 else 0


def wavelet_denoise(vol_data, vol, alpha, levels=None):
    """lib.rs:430-465 by explicit matrices: analysis down `levels` levels, every
    detail coefficient and the deepest approximation shrunk, synthesis up, each
    level's output cut to the next finer approximation's size."""
    L = wavelet_levels(vol) if levels is None else levels
    x = vol_data.reshape(vol[2], vol[1], vol[0]).astype(np.complex128)
    if L <= 0:
        return shrink(x, alpha).ravel()
    boxes, dims = [], []
    cur = x
    for _ in range(L):
        n = (cur.shape[2], cur.shape[1], cur.shape[0])
        c = apply3(cur, [analysis_matrix(n[a]) for a in range(3)])
        boxes.append(c)
        dims.append(n)
        h = [c.shape[2] // 2, c.shape[1] // 2, c.shape[0] // 2]
        cur = c[:h[2], :h[1], :h[0]]
    for l, c in enumerate(boxes):
        h = [c.shape[2] // 2, c.shape[1] // 2, c.shape[0] // 2]
        keep = c[:h[2], :h[1], :h[0]].copy()
        boxes[l] = shrink(c, alpha)
        if l != L - 1:
            boxes[l][:h[2], :h[1], :h[0]] = keep
    rec = None
    for l in range(L - 1, -1, -1):
        c = boxes[l]
        h = [c.shape[2] // 2, c.shape[1] // 2, c.shape[0] // 2]
        if rec is not None:
            c = c.copy()
            c[:h[2], :h[1], :h[0]] = rec
        n = dims[l]
        rec = apply3(c, [synthesis_matrix(n[a], h[a]) for a in range(3)])
    return rec.ravel()


# ------------------------------------------------------------ the loop
def recon(y, vol, nq, blk, shifts, mode="svt", lam_factor=0.02, rank=0, alpha=0.0, S=None, order="llr_first",
          normalize=True, img_corr=None, x0=None, momentum=0.0, mom_prev="handed_over"):
    """lib.rs:224-341: energy, corrections, zero-filled adjoint, threshold from the
    adjoint's blocks at shift 0, then per iteration LLR -> wavelet -> forward ->
    hard projection (residual / total energy) -> adjoint.

    x0 replaces the first iterate after the threshold is taken. momentum > 0
    (the warm start's extension): the LLR step receives
    x_k + momentum (x_k - x_{k-1}); mom_prev="extrapolated" keeps the
    extrapolated iterate as x_{k-1} instead (a planted mistake)."""
    N = int(np.prod(vol))
    y = y.astype(np.complex128)
    energy = float(np.sum(np.abs(y) ** 2))
    kc = linear_corrections(y, vol, nq)
    ic = np.zeros((nq, 3)) if img_corr is None else np.asarray(img_corr, np.float64).reshape(nq, 3)
    x = signal_model(y, vol, nq, kc, ic, forward=False)
    lam = lam_factor * llr_step(x.copy(), vol, nq, blk, (0, 0, 0), S, mode, 0.0, rank, modify=False)
    if x0 is not None:
        x = np.asarray(x0, np.complex128).copy()
    hist = []
    prev = None
    for it in range(len(shifts)):
        if momentum > 0:
            cur = x.copy()
            if prev is not None:
                x[:] = cur + momentum * (cur - prev)
            prev = x.copy() if mom_prev == "extrapolated" else cur

        def llr():
            llr_step(x, vol, nq, blk, tuple(int(s) for s in shifts[it]), S, mode, lam, rank)

        def wav():
            if alpha > 0:
                for q in range(nq):
                    x[q * N:(q + 1) * N] = wavelet_denoise(x[q * N:(q + 1) * N], vol, alpha)

        for step in ((llr, wav) if order == "llr_first" else (wav, llr)):
            step()
        k = signal_model(x, vol, nq, kc, ic, forward=True)
        meas = y != 0
        err = float(np.sum(np.abs(k[meas] - y[meas]) ** 2))  # hard_project, lib.rs:392-406
        k[meas] = y[meas]
        hist.append(err / energy if normalize else err)
        x[:] = signal_model(k, vol, nq, kc, ic, forward=False)
    return x, np.array(hist), lam, kc


def measurements(vol, nq, frac, seed):
    r = np.random.default_rng(seed)
    N = int(np.prod(vol))
    x = (r.standard_normal(N * nq) + 1j * r.standard_normal(N * nq)) * (1.0 + 2.0 * (r.random(N * nq) < 0.2))
    z = np.zeros((nq, 3))
    y = signal_model(x, vol, nq, z, z, forward=True)
    y[r.random(y.size) >= frac] = 0
    return y


def oracle_recon(y, vol, nq, blk, shifts, mode="svt", lam_factor=0.02, rank=0, alpha=0.0, S=None, img_corr=None):
    return O.recon(y.astype(np.complex128), vol, nq, blk, shifts, llr_mode={"svt": 1, "rank": 0}[mode], rank=rank,
                   lambda_factor=lam_factor, wavelet_alpha=alpha, subspace=S, img_corr=img_corr)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


CASES = [
    # vol, nq, blk, svt factor, wavelet alpha, subspace rank
    ((12, 8, 8), 3, (4, 4, 4), 0.3, 0.4, 0),
    ((12, 12, 12), 3, (4, 4, 4), 0.3, 0.4, 0),
    ((8, 8, 8), 4, (4, 4, 2), 0.3, 0.3, 2),
]


@pytest.mark.parametrize("vol,nq,blk,lf,alpha,srank", CASES)
def test_loop_matches_independent_restatement(vol, nq, blk, lf, alpha, srank):
    y = measurements(vol, nq, 0.5, 7)
    shifts = np.array([[3, 1, 2], [5, 6, 7]]) % np.array(vol)
    S = None
    if srank:
        S = np.linalg.qr(crandn(nq, srank))[0]
    x_np, h_np, lam_np, kc_np = recon(y, vol, nq, blk, shifts, lam_factor=lf, alpha=alpha, S=S)
    x_o, h_o, lam_o, kc_o = oracle_recon(y, vol, nq, blk, shifts, lam_factor=lf, alpha=alpha, S=S)
    assert np.array_equal(kc_o, kc_np.astype(np.float32))
    assert abs(lam_o - lam_np) <= 1e-10 * lam_np
    assert np.allclose(h_o, h_np, rtol=1e-9, atol=0)
    assert rel(x_o, x_np) <= 1e-9
    # the checks above must fail for a loop with LLR and wavelet swapped, or a
    # residual not divided by the total energy
    x_sw, h_sw, _, _ = recon(y, vol, nq, blk, shifts, lam_factor=lf, alpha=alpha, S=S, order="wavelet_first")
    assert rel(x_o, x_sw) > 1e-4 and not np.allclose(h_o, h_sw, rtol=1e-6)
    _, h_raw, _, _ = recon(y, vol, nq, blk, shifts, lam_factor=lf, alpha=alpha, S=S, normalize=False)
    assert not np.allclose(h_o, h_raw, rtol=1e-3)


def test_loop_with_first_iterate_and_momentum_matches_independent_restatement():
    """The oracle's momentum (image space, first iteration only records x_0)
    and given first iterate against the numpy restatement, three iterations;
    the bar must reject momentum that keeps the extrapolated iterate as the
    previous one, and the plain loop from the same start."""
    vol, nq, blk, S = (8, 8, 8), 4, (4, 4, 2), np.linalg.qr(crandn(4, 2))[0]
    y = measurements(vol, nq, 0.5, 11)
    shifts = np.array([[3, 1, 2], [5, 6, 7], [1, 0, 4]]) % np.array(vol)
    x0 = crandn(int(np.prod(vol)) * nq)
    kw = dict(lam_factor=0.3, alpha=0.3, S=S)
    x_np, h_np, lam_np, _ = recon(y, vol, nq, blk, shifts, x0=x0, momentum=0.6, **kw)
    x_o, h_o, lam_o, _ = O.recon(y.astype(np.complex128), vol, nq, blk, shifts, llr_mode=1, lambda_factor=0.3,
                                 wavelet_alpha=0.3, subspace=S, x0=x0, momentum=0.6)
    assert abs(lam_o - lam_np) <= 1e-10 * lam_np
    assert np.allclose(h_o, h_np, rtol=1e-9, atol=0)
    assert rel(x_o, x_np) <= 1e-9
    x_bad, _, _, _ = recon(y, vol, nq, blk, shifts, x0=x0, momentum=0.6, mom_prev="extrapolated", **kw)
    x_plain, _, _, _ = recon(y, vol, nq, blk, shifts, x0=x0, **kw)
    assert rel(x_o, x_bad) > 1e-4 and rel(x_o, x_plain) > 1e-4
    # momentum 0 with a given start is the plain loop from that start
    x_o0, _, _, _ = O.recon(y.astype(np.complex128), vol, nq, blk, shifts, llr_mode=1, lambda_factor=0.3,
                            wavelet_alpha=0.3, subspace=S, x0=x0)
    assert rel(x_o0, x_plain) <= 1e-9


def test_loop_rank_mode_matches_independent_restatement():
    vol, nq, blk = (8, 8, 8), 4, (4, 4, 4)
    y = measurements(vol, nq, 0.4, 9)
    shifts = np.array([[1, 2, 3], [7, 0, 4]])
    x_np, h_np, _, _ = recon(y, vol, nq, blk, shifts, mode="rank", rank=2)
    x_o, h_o, _, _ = oracle_recon(y, vol, nq, blk, shifts, mode="rank", rank=2)
    assert np.allclose(h_o, h_np, rtol=1e-9, atol=0)
    assert rel(x_o, x_np) <= 1e-9


def test_loop_with_image_corrections_matches_independent_restatement():
    """Sub-voxel image corrections (lin_img_corr of signal_model.rs:34-79, fixed at
    zero by lib.rs:148 but an input of the signal model) in both directions."""
    vol, nq, blk = (8, 12, 8), 3, (4, 4, 4)
    y = measurements(vol, nq, 0.5, 17)
    ic = np.array([[0.3, -1.25, 2.0], [0.0, 0.5, -0.75], [1.5, 0.0, 0.25]], np.float32)
    shifts = np.array([[1, 2, 3], [6, 5, 4]])
    x_np, h_np, _, _ = recon(y, vol, nq, blk, shifts, lam_factor=0.3, alpha=0.4, img_corr=ic)
    x_o, h_o, _, _ = oracle_recon(y, vol, nq, blk, shifts, lam_factor=0.3, alpha=0.4, img_corr=ic)
    x_0, _, _, _ = oracle_recon(y, vol, nq, blk, shifts, lam_factor=0.3, alpha=0.4)
    assert np.allclose(h_o, h_np, rtol=1e-9, atol=0)
    assert rel(x_o, x_np) <= 1e-9
    assert rel(x_o, x_0) > 1e-3


def test_fully_sampled_loop_returns_the_adjoint():
    """Every sample measured: the hard projection replaces all of k-space by y,
    so after every iteration the image is A^H y whatever the LLR and wavelet
    steps did (lib.rs:311-333 order: forward, project, adjoint)."""
    vol, nq, blk = (8, 12, 8), 3, (4, 4, 4)
    y = measurements(vol, nq, 1.01, 11)
    assert np.all(y != 0)
    kc = linear_corrections(y, vol, nq)
    adj = signal_model(y, vol, nq, kc, np.zeros((nq, 3)), forward=False)
    for n in (1, 3):
        x_o, h_o, _, _ = oracle_recon(y, vol, nq, blk, np.array([[1, 2, 3]] * n), lam_factor=0.5, alpha=0.5)
        assert rel(x_o, adj) <= 1e-12
        assert np.all(h_o > 1e-3)  # the prox steps did move the iterate


def test_identity_prox_keeps_the_zero_filled_adjoint():
    """Rank truncation to all columns and no wavelet is the identity: A x0 = y on
    the measured samples (A unitary, x0 = A^H y), so x0 is a fixed point and
    the residual is 0."""
    vol, nq, blk = (8, 8, 8), 3, (4, 4, 4)
    y = measurements(vol, nq, 0.3, 13)
    kc = linear_corrections(y, vol, nq)
    x0 = signal_model(y, vol, nq, kc, np.zeros((nq, 3)), forward=False)
    x_o, h_o, _, _ = oracle_recon(y, vol, nq, blk, np.array([[1, 2, 3], [4, 5, 6]]), mode="rank", rank=nq)
    assert rel(x_o, x0) <= 1e-12
    assert np.all(h_o <= 1e-24)


# ------------------------------------------------------------ wavelet
def test_explicit_single_level_matrices_reconstruct():
    for n in (4, 5, 8, 11, 12):
        W = analysis_matrix(n)
        Sm = synthesis_matrix(n, W.shape[0] // 2)
        assert np.allclose(Sm @ W, np.eye(n), atol=1e-12)


@pytest.mark.parametrize("vol", [(12, 12, 12), (13, 12, 11), (24, 20, 14), (7, 9, 6)])
@pytest.mark.parametrize("alpha", [0.0, 0.6])
def test_multilevel_wavelet_matches_explicit_matrices(vol, alpha):
    x = crandn(int(np.prod(vol)))
    ref = wavelet_denoise(x, vol, alpha)
    got = x.copy()
    O.wavelet_denoise(got, vol, alpha)
    assert rel(got, ref) <= 1e-12
    if alpha == 0.0:
        assert rel(got, x) <= 1e-12


def test_single_level_coefficients_match_explicit_matrices():
    vol = (13, 12, 11)
    x = crandn(int(np.prod(vol)))
    coef, cdims = O.dwt3(x.copy(), vol)
    ref = apply3(x.reshape(vol[2], vol[1], vol[0]), [analysis_matrix(vol[a]) for a in range(3)])
    assert list(ref.shape[::-1]) == list(cdims)
    assert rel(coef, ref.ravel()) <= 1e-12


def test_wavelet_shrink_set_and_level_rule_are_what_the_oracle_does():
    """The explicit construction with another level count, or with the deepest
    approximation left unshrunk, differs from the oracle: these checks see both."""
    vol, alpha = (24, 20, 14), 0.6
    x = crandn(int(np.prod(vol)))
    got = x.copy()
    O.wavelet_denoise(got, vol, alpha)
    assert wavelet_levels(vol) == 2
    assert rel(got, wavelet_denoise(x, vol, alpha, levels=1)) > 1e-3
    assert rel(got, wavelet_denoise(x, vol, alpha, levels=3)) > 1e-3
