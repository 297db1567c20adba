"""The reconstruction loop of llr_recon_exec (lib.rs:113-363) on one GPU, plus
the lib.rs / data_import.rs helpers it calls.

`Recon` keeps the iterate, the measurements and every workspace in HBM for
the whole run (llr_recon_* in include/llr_hip.h); only the per-iteration
residuals come back to the host.
"""
import ctypes
import sys

import numpy as np

from . import _lib

_sz = ctypes.c_size_t


def hard_project(x, y, ctx=None):
    """lib.rs:392-406: where y != 0, x = y; returns sum |x - y|^2 (f64). In place on x."""
    X = _lib.as_cf32(x, "x")
    Y = _lib.as_cf32(y, "y")
    if X.size != Y.size:
        raise ValueError("dataset and measurements must have the same dimensions")
    err = ctypes.c_double()
    _lib.check(_lib.load().llr_hard_project(_lib.ctx_handle(ctx), _lib.ptr(X), _lib.ptr(Y), _sz(X.size),
                                            ctypes.byref(err)))
    return err.value


def calc_undersampling_factor(y, ctx=None):
    """lib.rs:409-417: (len / nonzero, total energy)."""
    Y = _lib.as_cf32(y, "y")
    f = ctypes.c_double()
    e = ctypes.c_double()
    _lib.check(_lib.load().llr_calc_undersampling_factor(_lib.ctx_handle(ctx), _lib.ptr(Y), _sz(Y.size),
                                                         ctypes.byref(f), ctypes.byref(e)))
    return f.value, e.value


def under_sample_kspace(data, dims, nq, mask, ctx=None):
    """data_import.rs:416-438: zero readout lines whose (y, z, q) mask entry is 0. In place."""
    D = _lib.as_cf32(data, "data")
    M = _lib.as_cf32(mask, "mask")
    if M.size != dims[1] * dims[2] * nq:
        raise ValueError("phase encoding dimension mismatch")
    _lib.check(_lib.load().llr_under_sample_kspace(_lib.ctx_handle(ctx), _lib.ptr(D), _lib.sz3(dims), _sz(nq),
                                                   _lib.ptr(M)))


def wavelet_denoise(volumes, n_volumes, vol_size, alpha, levels=-1, ctx=None):
    """lib.rs:430-458 for n_volumes volumes, in place."""
    V = _lib.as_cf32(volumes, "volumes")
    _lib.check(_lib.load().llr_wavelet_denoise(_lib.ctx_handle(ctx), _lib.ptr(V), n_volumes, _lib.sz3(vol_size),
                                               float(alpha), int(levels)))


def make_params(vol, nq, shifts, blk=(8, 8, 8), llr_mode="svt", rank=0, svt_lambda=0.0, lambda_factor=0.02,
                wavelet_alpha=0.0, wavelet_levels=-1, subspace=None, residual_stop=0.0, precision="fp32",
                partial_blocks=False, storage="fp32", hybrid=True, img_corr=None, fused_projection=None):
    """llr_recon_params for the C ABI, plus the arrays it points at (keep them alive)."""
    vol = [int(v) for v in vol]
    nq = int(nq)
    sh = np.ascontiguousarray(np.asarray(shifts, np.uint64).reshape(-1, 3))
    p = _lib.ReconParams()
    p.vol[:] = vol
    p.nq = nq
    p.blk[:] = [int(b) for b in blk]
    p.n_iter = sh.shape[0]
    p.llr_mode = {"rank": _lib.LLR_MODE_RANK, "svt": _lib.LLR_MODE_SVT}[llr_mode]
    p.precision = {"fp32": _lib.LLR_PREC_FP32, "mixed": _lib.LLR_PREC_MIXED}[precision]
    p.rank = int(rank)
    p.svt_lambda = float(svt_lambda)
    p.lambda_factor = float(lambda_factor)
    p.wavelet_alpha = float(wavelet_alpha)
    p.wavelet_levels = int(wavelet_levels)
    sub = None
    if subspace is not None:
        sub = np.asfortranarray(np.asarray(subspace, np.complex64))
        if sub.shape[0] != nq:
            raise ValueError("unexpected subspace dimension")
        p.srank = sub.shape[1]
        sub = sub.ravel(order="F").copy()
        p.subspace = ctypes.c_void_p(sub.ctypes.data)
    else:
        p.srank = 0
        p.subspace = None
    p.shifts = ctypes.c_void_p(sh.ctypes.data) if sh.shape[0] else None
    p.residual_stop = float(residual_stop)
    p.partial_blocks = 1 if partial_blocks else 0
    p.storage = {"fp32": _lib.LLR_STORE_FP32, "fp16": _lib.LLR_STORE_FP16}[storage]
    p.no_hybrid = 0 if hybrid else 1
    p.fused_projection = 0 if fused_projection is None else (1 if fused_projection else -1)
    ic = None
    if img_corr is not None:
        ic = np.ascontiguousarray(np.asarray(img_corr, np.float32).reshape(-1))
        if ic.size != 3 * nq:
            raise ValueError("the number of linear corrections must be equal to the number of volumes")
        p.lin_img_corr = ctypes.c_void_p(ic.ctypes.data)
    return p, (sh, sub, ic)


class Recon:
    """Device-resident LLR reconstruction (llr_recon_exec's loop, lib.rs:227-341).

    llr_mode: "svt" (singular-value soft threshold, lambda = lambda_factor *
    sigma_max of the starting blocks unless svt_lambda > 0) or "rank" (hard
    truncation to `rank`, low_rank.rs:10-69).
    partial_blocks: block grid ceil(vol/blk) with edge blocks truncated at the
    volume end, instead of floor(vol/blk) (block.rs:112-118).
    storage: "fp32" (complex64) or "fp16" (two IEEE half-precision numbers per
    complex value for the iterate, the measurements' working copy and the
    wavelet workspace, held at a scale of sqrt(voxels / measurement energy);
    precision "mixed" with a subspace only). Arithmetic is FP32 either way and
    every output is in the caller's units.
    hybrid: when the measurements are a phase-encode mask (every x-line fully
    measured or empty), iterate in (x, ky, kz) without the x transforms, which
    commute with the hard projection (same result up to rounding); False keeps
    the full 3D forward model and adjoint in every iteration.
    img_corr: (nq, 3) linear image corrections of the signal model (a
    sub-voxel image shift per volume; signal_model.rs:34-79), default zero.
    fused_projection: with the hybrid iteration and the subspace wavelet fused
    into level 0, the adjoint's inverse y pass can hand the next LLR step the
    subspace coefficients directly, the image being formed once at the end of
    a run (same result up to rounding). None: with complex64 storage only
    (where it is faster); True: with either storage; False: never (the image is
    stored and projected in every iteration).
    """

    def __init__(self, vol, nq, shifts, blk=(8, 8, 8), llr_mode="svt", rank=0, svt_lambda=0.0,
                 lambda_factor=0.02, wavelet_alpha=0.0, wavelet_levels=-1, subspace=None, residual_stop=0.0,
                 precision="fp32", partial_blocks=False, storage="fp32", hybrid=True, img_corr=None, ctx=None,
                 fused_projection=None):
        self.ctx = ctx or _lib.default_context()
        p, self._keep = make_params(vol, nq, shifts, blk, llr_mode, rank, svt_lambda, lambda_factor, wavelet_alpha,
                                    wavelet_levels, subspace, residual_stop, precision, partial_blocks, storage,
                                    hybrid, img_corr, fused_projection)
        self.vol = [int(v) for v in vol]
        self.nq = int(nq)
        self.D = int(np.prod(self.vol)) * self.nq
        self.n_iter = int(p.n_iter)
        self.srank = int(p.srank)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().llr_recon_create(self.ctx.handle, ctypes.byref(p), ctypes.byref(h)))
        self.handle = h

    def set_measurements(self, y=None, device_ptr=None):
        if device_ptr is not None:
            _lib.check(_lib.load().llr_recon_set_measurements(self.handle, ctypes.c_void_p(int(device_ptr)),
                                                              _sz(self.D), 1))
        else:
            Y = _lib.as_cf32(y, "measurements")
            if Y.size != self.D:
                raise ValueError("measurements must hold nx*ny*nz*nq samples")
            _lib.check(_lib.load().llr_recon_set_measurements(self.handle, _lib.ptr(Y), _sz(self.D), 0))

    def set_initial(self, x=None, device_ptr=None):
        if device_ptr is not None:
            _lib.check(_lib.load().llr_recon_set_initial(self.handle, ctypes.c_void_p(int(device_ptr)), _sz(self.D), 1))
        else:
            X = _lib.as_cf32(x, "initial image")
            _lib.check(_lib.load().llr_recon_set_initial(self.handle, _lib.ptr(X), _sz(self.D), 0))

    def prepare(self):
        """Set-up without iterating: statistics, corrections, zero-filled adjoint
        (or the set initial image) and the SVT threshold; the next run() starts here."""
        _lib.check(_lib.load().llr_recon_prepare(self.handle))

    def coefficients(self):
        """Subspace coefficients C = X S of the current iterate, (srank, nz, ny, nx) complex64 flat planes."""
        out = np.empty(self.srank * (self.D // self.nq), np.complex64)
        _lib.check(_lib.load().llr_recon_get_coefficients(self.handle, _lib.ptr(out), _sz(out.size), 0))
        return out

    def coefficients_to_device(self, device_ptr, n):
        _lib.check(_lib.load().llr_recon_get_coefficients(self.handle, ctypes.c_void_p(int(device_ptr)), _sz(n), 1))

    def set_coefficients(self, c=None, device_ptr=None, n=None):
        if device_ptr is not None:
            _lib.check(_lib.load().llr_recon_set_coefficients(self.handle, ctypes.c_void_p(int(device_ptr)), _sz(n), 1))
        else:
            C = _lib.as_cf32(c, "coefficients")
            _lib.check(_lib.load().llr_recon_set_coefficients(self.handle, _lib.ptr(C), _sz(C.size), 0))

    def warm_start(self, net, to_net=None, from_net=None):
        """Iterate <- the U-Net's prediction from the zero-filled adjoint's
        coefficients, made data-consistent. Returns (input scale, relative
        data-consistency error of the prediction)."""
        from .unet import _mix, _mptr
        K = self.srank
        tn, fn = _mix(to_net, K), _mix(from_net, K)
        sc = ctypes.c_double()
        res = ctypes.c_double()
        _lib.check(_lib.load().llr_recon_warm_start(self.handle, net.handle, _mptr(tn), _mptr(fn), ctypes.byref(sc),
                                                    ctypes.byref(res)))
        return sc.value, res.value

    def run_learned(self, net, residual_stop, to_net=None, from_net=None, max_iter=0, momentum=0.0, min_iter=0):
        """Warm start + iterations (at least min_iter, with momentum in coefficient space
        when momentum > 0) to residual_stop, cold-start fallback (llr_recon_run_learned).
        Returns (relative residual history, fell_back)."""
        from .unet import _mix, _mptr
        K = self.srank
        tn, fn = _mix(to_net, K), _mix(from_net, K)
        wp = _lib.WarmParams()
        wp.to_net = ctypes.c_void_p(tn.ctypes.data) if tn is not None else None
        wp.from_net = ctypes.c_void_p(fn.ctypes.data) if fn is not None else None
        wp.residual_stop = float(residual_stop)
        wp.max_iter = int(max_iter)
        wp.momentum = float(momentum)
        wp.min_iter = int(min_iter)
        hist = np.zeros(max(1, self.n_iter), np.float64)
        done = ctypes.c_size_t()
        fb = ctypes.c_int()
        _lib.check(_lib.load().llr_recon_run_learned(self.handle, net.handle, ctypes.byref(wp), _lib.ptr(hist),
                                                     ctypes.byref(done), ctypes.byref(fb)))
        return hist[:done.value], bool(fb.value)

    def run(self):
        """Run the iterations; returns the relative residual history (lib.rs:321)."""
        hist = np.zeros(max(1, self.n_iter), np.float64)
        done = ctypes.c_size_t()
        _lib.check(_lib.load().llr_recon_run(self.handle, _lib.ptr(hist), ctypes.byref(done)))
        return hist[:done.value]

    def image(self):
        out = np.empty(self.D, np.complex64)
        _lib.check(_lib.load().llr_recon_get_image(self.handle, _lib.ptr(out), _sz(self.D), 0))
        return out

    def image_to_device(self, device_ptr):
        _lib.check(_lib.load().llr_recon_get_image(self.handle, ctypes.c_void_p(int(device_ptr)), _sz(self.D), 1))

    def info(self):
        lam = ctypes.c_double()
        kc = np.zeros(3 * self.nq, np.float32)
        _lib.check(_lib.load().llr_recon_info(self.handle, ctypes.byref(lam), _lib.ptr(kc)))
        return lam.value, kc.reshape(self.nq, 3)

    def hybrid(self):
        """True when the last set-up chose the (x, ky, kz) iteration."""
        h = ctypes.c_int()
        _lib.check(_lib.load().llr_recon_hybrid(self.handle, ctypes.byref(h)))
        return bool(h.value)

    def mixed_stats(self):
        """(blocks whose eigen-solve fell back to FP64, blocks processed) of the mixed form."""
        f = ctypes.c_ulonglong()
        n = ctypes.c_ulonglong()
        _lib.check(_lib.load().llr_recon_mixed_stats(self.handle, ctypes.byref(f), ctypes.byref(n)))
        return f.value, n.value

    def enable_timing(self, on=True):
        _lib.check(_lib.load().llr_recon_enable_timing(self.handle, int(on)))

    def stage_times(self):
        ms = np.zeros(_lib.LLR_NUM_STAGES, np.float64)
        n = np.zeros(_lib.LLR_NUM_STAGES, np.uint64)
        _lib.check(_lib.load().llr_recon_stage_times(self.handle, _lib.ptr(ms), _lib.ptr(n)))
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(_lib.STAGE_NAMES)}

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().llr_recon_destroy(self.handle)
            self.handle = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the HIP runtime may already be torn down: leave
        # the handle to the process exit instead of calling into it
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class ReconBatch:
    """Many reconstructions of uneven size in one graph-replayed loop
    (llr_recon_batch_*): `members` is a list of dicts of Recon keyword
    arguments (vol, nq, shifts, blk, subspace, ...); all share n_iter.
    autotune picks each member's block size from {6,8,10}^3 and launch shapes."""

    def __init__(self, members, autotune=False, ctx=None):
        self.ctx = ctx or _lib.default_context()
        ps, self._keep = [], []
        for m in members:
            p, keep = make_params(**m)
            ps.append(p)
            self._keep.append(keep)
        arr = (_lib.ReconParams * len(ps))(*ps)
        self.members = [dict(m) for m in members]
        self.D = [int(np.prod(m["vol"])) * int(m["nq"]) for m in members]
        self.n_iter = int(ps[0].n_iter)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().llr_recon_batch_create(self.ctx.handle, arr, _sz(len(ps)), int(bool(autotune)),
                                                      ctypes.byref(h)))
        self.handle = h

    def __len__(self):
        return len(self.D)

    def set_measurements(self, i, y=None, device_ptr=None):
        if device_ptr is not None:
            _lib.check(_lib.load().llr_recon_batch_set_measurements(self.handle, _sz(i), ctypes.c_void_p(int(device_ptr)),
                                                                    _sz(self.D[i]), 1))
        else:
            Y = _lib.as_cf32(y, "measurements")
            if Y.size != self.D[i]:
                raise ValueError("measurements must hold nx*ny*nz*nq samples")
            _lib.check(_lib.load().llr_recon_batch_set_measurements(self.handle, _sz(i), _lib.ptr(Y), _sz(Y.size), 0))

    def run(self):
        """Runs every member; returns the (n_members, n_iter) relative residual history."""
        hist = np.zeros((len(self), max(self.n_iter, 1)), np.float64)
        _lib.check(_lib.load().llr_recon_batch_run(self.handle, _lib.ptr(hist)))
        return hist[:, :self.n_iter]

    def image(self, i):
        out = np.empty(self.D[i], np.complex64)
        _lib.check(_lib.load().llr_recon_batch_get_image(self.handle, _sz(i), _lib.ptr(out), _sz(out.size), 0))
        return out

    def tuning(self, i):
        blk = (ctypes.c_size_t * 3)()
        tpw = (ctypes.c_int * 3)()
        stage = ctypes.c_int()
        lam = ctypes.c_double()
        _lib.check(_lib.load().llr_recon_batch_tuning(self.handle, _sz(i), blk, tpw, ctypes.byref(stage),
                                                      ctypes.byref(lam)))
        return dict(blk=tuple(blk), fft_tpw=tuple(tpw), llr_stage=stage.value, lam=lam.value)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().llr_recon_batch_destroy(self.handle)
            self.handle = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the HIP runtime may already be torn down: leave
        # the handle to the process exit instead of calling into it
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def wavelet_denoise_shifted(volumes, n_volumes, vol_size, alpha, shift, levels=-1, ctx=None):
    """Cycle-spinning wavelet shrinkage: wavelet_denoise of the volumes moved
    circularly by shift (sample i reads i + shift), moved back. In place."""
    V = _lib.as_cf32(volumes, "volumes")
    sh = (ctypes.c_int * 3)(*[int(s) for s in shift])
    _lib.check(_lib.load().llr_wavelet_denoise_shifted(_lib.ctx_handle(ctx), _lib.ptr(V), _sz(n_volumes),
                                                       _lib.sz3(vol_size), ctypes.c_float(alpha), int(levels), sh))
