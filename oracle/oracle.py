"""ctypes front end of the CPU oracle (liboracle.so, built from llr_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py as the checker. The product never imports this module.

Arrays are numpy, column-major in the reference's sense: a 4-D data set
(nx, ny, nz, nq) is stored flat with x fastest and the volume index slowest
(lib.rs:374-388), i.e. ``arr.ravel(order="F")``. Complex precision follows the
dtype: complex64 runs the FP32 reference path, complex128 the FP64 oracle.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_size_p = ctypes.POINTER(ctypes.c_size_t)


def build(force=False):
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"], check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        _LIB = ctypes.CDLL(so)
        _LIB.orc_hard_project_f32.restype = ctypes.c_double
        _LIB.orc_hard_project_f64.restype = ctypes.c_double
        _LIB.orc_low_rank_approx_batch_f32.restype = ctypes.c_double
        _LIB.orc_low_rank_approx_batch_f64.restype = ctypes.c_double
        _LIB.orc_svt_batch_f32.restype = ctypes.c_double
        _LIB.orc_svt_batch_f64.restype = ctypes.c_double
        _LIB.orc_svt_batch_f32.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_double,
                                           ctypes.c_size_t, ctypes.c_void_p]
        _LIB.orc_svt_batch_f64.argtypes = _LIB.orc_svt_batch_f32.argtypes
        _LIB.orc_wavelet_denoise_f32.argtypes = [ctypes.c_void_p, c_size_p, ctypes.c_double, ctypes.c_int]
        _LIB.orc_wavelet_denoise_f64.argtypes = _LIB.orc_wavelet_denoise_f32.argtypes
        _LIB.orc_coord_to_col_maj_index.restype = ctypes.c_size_t
    return _LIB


def threads(n=0):
    """Set (n > 0) and return the OpenMP thread count of the oracle."""
    return lib().orc_num_threads(int(n))


def _sz(v):
    a = (ctypes.c_size_t * len(v))(*[int(x) for x in v])
    return a


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _suf(a):
    if a.dtype == np.complex64:
        return "f32"
    if a.dtype == np.complex128:
        return "f64"
    raise TypeError(f"unsupported dtype {a.dtype}")


def _check(a):
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]


# ----------------------------------------------------------------- block.rs
def grid_dim(vol, blk):
    g = (ctypes.c_size_t * 3)()
    lib().orc_grid_dim(_sz(vol), _sz(blk), g)
    return [g[0], g[1], g[2]]


def lift_data_set(vols, nvols, mats, vol, blk, shift, start, n):
    """vols: flat array of nvols volumes; mats: flat output (any dtype)."""
    r = lib().orc_lift_data_set(_p(vols), ctypes.c_size_t(nvols), ctypes.c_size_t(vols.itemsize), _p(mats),
                                _sz(vol), _sz(blk), _sz(shift), ctypes.c_size_t(start), ctypes.c_size_t(n))
    if r:
        raise ValueError("block dimension must be less than or equal to volume dimension")


def un_lift_data_set(vols, nvols, mats, vol, blk, shift, start, n):
    r = lib().orc_un_lift_data_set(_p(vols), ctypes.c_size_t(nvols), ctypes.c_size_t(vols.itemsize), _p(mats),
                                   _sz(vol), _sz(blk), _sz(shift), ctypes.c_size_t(start), ctypes.c_size_t(n))
    if r:
        raise ValueError("block dimension must be less than or equal to volume dimension")


# ----------------------------------------------------- dti_subspace / blas
def cgemm(m, k, n, batch, broadcast_a, alpha, beta, ta, tb, A, B, Cm):
    s = _suf(Cm)
    rt = np.float32 if s == "f32" else np.float64
    al = np.array([alpha.real, alpha.imag], rt)
    be = np.array([beta.real, beta.imag], rt)
    getattr(lib(), "orc_cgemm_" + s)(ctypes.c_size_t(m), ctypes.c_size_t(k), ctypes.c_size_t(n),
                                     ctypes.c_size_t(batch), ctypes.c_int(int(broadcast_a)), _p(al), _p(be),
                                     ctypes.c_int(int(ta)), ctypes.c_int(int(tb)), _p(A), _p(B), _p(Cm))


def conj_transpose_matrix(data, m, n, batch):
    getattr(lib(), "orc_conj_transpose_" + _suf(data))(_p(data), ctypes.c_size_t(m), ctypes.c_size_t(n),
                                                      ctypes.c_size_t(batch))


def dti_project(S, X, Cm, batch, nq, nvox, srank):
    getattr(lib(), "orc_dti_project_" + _suf(X))(_p(S), _p(X), _p(Cm), *[ctypes.c_size_t(v) for v in
                                                                         (batch, nq, nvox, srank)])


def dti_back_project(S, X, Cm, batch, nq, nvox, srank):
    getattr(lib(), "orc_dti_back_project_" + _suf(X))(_p(S), _p(X), _p(Cm), *[ctypes.c_size_t(v) for v in
                                                                              (batch, nq, nvox, srank)])


def subspace_constraint(S, X, batch, nq, nvox, srank):
    getattr(lib(), "orc_subspace_constraint_" + _suf(X))(_p(S), _p(X), *[ctypes.c_size_t(v) for v in
                                                                         (batch, nq, nvox, srank)])


def low_rank_approx_batch(m, n, rank, batch, data):
    return getattr(lib(), "orc_low_rank_approx_batch_" + _suf(data))(
        ctypes.c_size_t(m), ctypes.c_size_t(n), ctypes.c_size_t(rank), ctypes.c_size_t(batch), _p(data))


def svt_batch(m, n, lam, batch, data):
    return getattr(lib(), "orc_svt_batch_" + _suf(data))(m, n, float(lam), batch, _p(data))


def llr_partial(x, vol, nq, blk, shift, mode=1, rank=0, lam=0.0, subspace=None):
    """One LLR step over the ceil(vol/blk) grid with truncated edge blocks, in place on x
    (nq volumes, column-major). Returns the largest singular value seen."""
    s = _suf(x)
    v = (ctypes.c_size_t * 3)(*vol)
    b = (ctypes.c_size_t * 3)(*blk)
    sh = (ctypes.c_size_t * 3)(*shift)
    sub, srank = None, 0
    if subspace is not None:
        sub = np.asfortranarray(subspace).astype(x.dtype).ravel(order="F").copy()
        srank = subspace.shape[1]
    f = getattr(lib(), "orc_llr_partial_" + s)
    f.restype = ctypes.c_double
    return f(_p(x), ctypes.c_size_t(nq), v, b, sh, ctypes.c_size_t(srank),
             None if sub is None else _p(sub), ctypes.c_int(mode), ctypes.c_int(rank), ctypes.c_double(lam))


def block_svals(m, n, batch, data):
    s = _suf(data)
    out = np.zeros(n * batch, np.float32 if s == "f32" else np.float64)
    getattr(lib(), "orc_block_svals_" + s)(ctypes.c_size_t(m), ctypes.c_size_t(n), ctypes.c_size_t(batch),
                                          _p(data), _p(out))
    return out.reshape(batch, n)


def herm_eig(A):
    """Jacobi eigen-solve of a Hermitian matrix (numpy 2-D). Returns (w desc, V)."""
    n = A.shape[0]
    a = np.asfortranarray(A).copy(order="F")
    s = _suf(a)
    w = np.zeros(n, np.float32 if s == "f32" else np.float64)
    V = np.zeros((n, n), a.dtype, order="F")
    getattr(lib(), "orc_herm_eig_" + s)(ctypes.c_int(n), _p(a), _p(w), _p(V))
    return w, V


# ------------------------------------------------------------ signal model
def fft3_batch(data, dims, nvols, sign):
    getattr(lib(), "orc_fft3_batch_" + _suf(data))(_p(data), _sz(dims), ctypes.c_size_t(nvols), ctypes.c_int(sign))


def signal_model(x, dims, nvols, ksp_corr, img_corr, direction, phase_mode=None):
    """direction 0 = Forward, 1 = Inverse (signal_model.rs:14-26)."""
    s = _suf(x)
    if phase_mode is None:
        phase_mode = 0 if s == "f32" else 1
    kc = np.ascontiguousarray(ksp_corr, np.float32).reshape(-1)
    ic = np.ascontiguousarray(img_corr, np.float32).reshape(-1)
    assert kc.size == 3 * nvols and ic.size == 3 * nvols
    getattr(lib(), "orc_signal_model_" + s)(_p(x), _sz(dims), ctypes.c_size_t(nvols), _p(kc), _p(ic),
                                           ctypes.c_int(direction), ctypes.c_int(phase_mode))


def estimate_linear_correction(ksp, dims):
    out = np.zeros(3, np.float32)
    getattr(lib(), "orc_estimate_linear_correction_" + _suf(ksp))(_p(ksp), _sz(dims), _p(out))
    return out


# ----------------------------------------------------------------- wavelet
def wavelet_denoise(vol, dims, alpha, levels=-1):
    getattr(lib(), "orc_wavelet_denoise_" + _suf(vol))(_p(vol), _sz(dims), float(alpha), int(levels))


def dwt3(vol, dims):
    nc = [(d + 3) // 2 for d in dims]
    out = np.zeros(8 * nc[0] * nc[1] * nc[2], vol.dtype)
    getattr(lib(), "orc_dwt3_" + _suf(vol))(_p(vol), _sz(dims), _p(out))
    return out, [2 * c for c in nc]


def idwt3(coef, cdims, dims):
    out = np.zeros(int(np.prod(dims)), coef.dtype)
    getattr(lib(), "orc_idwt3_" + _suf(coef))(_p(coef), _sz(cdims), _p(out), _sz(dims))
    return out


# ------------------------------------------------------- data consistency
def hard_project(x, y):
    return getattr(lib(), "orc_hard_project_" + _suf(x))(_p(x), _p(y), ctypes.c_size_t(x.size))


def calc_undersampling_factor(y):
    us = ctypes.c_double()
    e = ctypes.c_double()
    getattr(lib(), "orc_calc_undersampling_" + _suf(y))(_p(y), ctypes.c_size_t(y.size), ctypes.byref(us),
                                                       ctypes.byref(e))
    return us.value, e.value


def under_sample_kspace(data, dims, nq, mask):
    getattr(lib(), "orc_under_sample_kspace_" + _suf(data))(_p(data), _sz(dims), ctypes.c_size_t(nq), _p(mask))


def fftshift(data, dims, inverse=False):
    lib().orc_fftshift(_p(data), ctypes.c_size_t(data.itemsize), _sz(dims), ctypes.c_int(int(inverse)))


def phase_shift3(x, dims, sign):
    lib().orc_phase_shift3_f32(_p(x), _sz(dims), ctypes.c_int(sign))


def coord_to_col_maj_index(coords, dims):
    c = (ctypes.c_longlong * len(coords))(*coords)
    return lib().orc_coord_to_col_maj_index(c, _sz(dims), ctypes.c_size_t(len(dims)))


# ------------------------------------------------------------------- recon
class ReconParams(ctypes.Structure):
    _fields_ = [
        ("vol", ctypes.c_size_t * 3),
        ("nq", ctypes.c_size_t),
        ("blk", ctypes.c_size_t * 3),
        ("n_iter", ctypes.c_size_t),
        ("llr_mode", ctypes.c_int),
        ("phase_mode", ctypes.c_int),
        ("rank", ctypes.c_size_t),
        ("svt_lambda", ctypes.c_double),
        ("lambda_factor", ctypes.c_double),
        ("wavelet_alpha", ctypes.c_double),
        ("wavelet_levels", ctypes.c_int),
        ("srank", ctypes.c_size_t),
        ("subspace", ctypes.c_void_p),
        ("shifts", ctypes.POINTER(ctypes.c_size_t)),
        ("partial_blocks", ctypes.c_int),
        ("snap_iter", ctypes.c_size_t),
        ("snap", ctypes.c_void_p),
        ("img_corr", ctypes.c_void_p),
        ("x0", ctypes.c_void_p),
        ("momentum", ctypes.c_double),
    ]


def recon(meas, vol, nq, blk, shifts, llr_mode=1, rank=0, svt_lambda=0.0, lambda_factor=0.02,
          wavelet_alpha=0.0, wavelet_levels=-1, subspace=None, phase_mode=None, partial_blocks=False,
          snapshot_after=0, img_corr=None, x0=None, momentum=0.0):
    """Run the reconstruction loop of llr_recon_exec (lib.rs:227-341) on CPU.

    partial_blocks: ceil(vol/blk) block grid with truncated edge blocks (an
    extension; the reference grid is floor(vol/blk), block.rs:112-118).

    meas: flat k-space measurements (complex64 -> FP32 path, complex128 -> FP64 oracle).
    shifts: (n_iter, 3) integer block shifts. Returns (image, err_hist, lambda, ksp_corr).
    snapshot_after > 0: also return the iterate after that many iterations
    (a fifth element).
    x0: the first iterate (flat, same dtype as meas) in place of the
    zero-filled adjoint; the threshold still comes from the adjoint.
    momentum > 0: the LLR step acts on x_k + momentum (x_k - x_{k-1}) (an
    extension with no counterpart in the reference; the first iteration only
    records x_0).
    """
    s = _suf(meas)
    if phase_mode is None:
        phase_mode = 0 if s == "f32" else 1
    shifts = np.ascontiguousarray(shifts, np.uint64).reshape(-1, 3)
    p = ReconParams()
    p.vol[:] = list(vol)
    p.nq = nq
    p.blk[:] = list(blk)
    p.n_iter = shifts.shape[0]
    p.llr_mode = llr_mode
    p.phase_mode = phase_mode
    p.rank = rank
    p.svt_lambda = svt_lambda
    p.lambda_factor = lambda_factor
    p.wavelet_alpha = wavelet_alpha
    p.wavelet_levels = wavelet_levels
    sub = None
    if subspace is not None:
        sub = np.asfortranarray(subspace).astype(meas.dtype).ravel(order="F").copy()
        p.srank = subspace.shape[1]
        p.subspace = ctypes.c_void_p(sub.ctypes.data)
    else:
        p.srank = 0
        p.subspace = None
    p.shifts = shifts.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t))
    p.partial_blocks = 1 if partial_blocks else 0
    ic = None
    if img_corr is not None:
        ic = np.ascontiguousarray(img_corr, np.float32).reshape(-1)
        assert ic.size == 3 * nq
        p.img_corr = ctypes.c_void_p(ic.ctypes.data)
    xs = None
    if x0 is not None:
        xs = np.ascontiguousarray(x0, meas.dtype).reshape(-1)
        assert xs.size == meas.size
        p.x0 = ctypes.c_void_p(xs.ctypes.data)
    p.momentum = float(momentum)
    snap = None
    if snapshot_after:
        snap = np.zeros_like(meas)
        p.snap_iter = int(snapshot_after)
        p.snap = ctypes.c_void_p(snap.ctypes.data)
    x = np.zeros_like(meas)
    hist = np.zeros(max(1, shifts.shape[0]), np.float64)
    lam = ctypes.c_double()
    kc = np.zeros(3 * nq, np.float32)
    r = getattr(lib(), "orc_recon_" + s)(ctypes.byref(p), _p(meas), _p(x), _p(hist), ctypes.byref(lam), _p(kc))
    if r:
        raise ValueError(f"oracle recon rejected the parameters (code {r})")
    if snap is not None:
        return x, hist[:shifts.shape[0]], lam.value, kc.reshape(nq, 3), snap
    return x, hist[:shifts.shape[0]], lam.value, kc.reshape(nq, 3)


# --------------------------------------------------- diffusion_models.rs (FP64)
def b0_mask(bvecs, tol=1e-6):
    """data_import.rs:322-328: a volume is a b0 when |g| <= tol."""
    g = np.asarray(bvecs, np.float64)
    return np.sqrt(np.sum(g * g, axis=1)) <= tol


def dti_design_matrix(bvals, bvecs):
    """build_design_matrix (diffusion_models.rs:221-253): rows = diffusion-weighted
    volumes, columns {Dxx, Dyy, Dzz, Dxy, Dxz, Dyz}: -b gx^2, -b gy^2, -b gz^2,
    -2b gx gy, -2b gx gz, -2b gy gz."""
    b = np.asarray(bvals, np.float64)
    g = np.asarray(bvecs, np.float64)
    return np.stack([-b * g[:, 0] ** 2, -b * g[:, 1] ** 2, -b * g[:, 2] ** 2, -2 * b * g[:, 0] * g[:, 1],
                     -2 * b * g[:, 0] * g[:, 2], -2 * b * g[:, 1] * g[:, 2]], axis=1)


def dti_scalars(vols, n_vox, bvals, bvecs, b0_tol=1e-6):
    """calc_scalars (diffusion_models.rs:21-129) in FP64 on a flat (nq x n_vox)
    complex data set: b0 average of |x| (:42-50), rhs = ln(clamp(|x|/s0, FLT_MIN, 1))
    with s0 = max(b0av, FLT_MIN) (:205-216), least-squares tensor (:191-202),
    eigh of the lower triangle (:160-186), e1 >= e2 >= e3, MD, RD and
    FA = sqrt(3/2) |l - MD| / |l| (:111-116; 0/0 gives NaN as there).
    Returns a (6, n_vox) float64 array: e1, e2, e3, FA, MD, RD."""
    x = np.abs(np.asarray(vols).reshape(-1, n_vox)).astype(np.float64)
    m = b0_mask(bvecs, b0_tol)
    tiny = float(np.finfo(np.float32).tiny)
    s0 = np.maximum(x[m].mean(axis=0), tiny)
    dw = np.flatnonzero(~m)
    A = dti_design_matrix(np.asarray(bvals)[dw], np.asarray(bvecs)[dw])
    rhs = np.log(np.clip(x[dw] / s0, tiny, 1.0))
    D, *_ = np.linalg.lstsq(A, rhs, rcond=None)  # (6, n_vox)
    T = np.zeros((n_vox, 3, 3))
    T[:, 0, 0], T[:, 1, 1], T[:, 2, 2] = D[0], D[1], D[2]
    T[:, 1, 0], T[:, 2, 0], T[:, 2, 1] = D[3], D[4], D[5]
    w = np.linalg.eigvalsh(T, UPLO="L")  # ascending
    e1, e2, e3 = w[:, 2], w[:, 1], w[:, 0]
    md = (e1 + e2 + e3) / 3
    rd = (e2 + e3) / 2
    with np.errstate(invalid="ignore", divide="ignore"):
        num = np.sqrt((e1 - md) ** 2 + (e2 - md) ** 2 + (e3 - md) ** 2)
        den = np.sqrt(e1 ** 2 + e2 ** 2 + e3 ** 2)
        fa = np.sqrt(1.5) * num / den
    return np.stack([e1, e2, e3, fa, md, rd])
