"""cudwt3.rs API on the GPU: one-level wavelet transforms along one axis with
any even-length filter pair and a circular shift hook (cudwt3.rs:102-351,
kernels.cu:155-349). Arrays are flat column-major complex64."""
import ctypes

import numpy as np

from . import _lib

_sz = ctypes.c_size_t

# db2 (pywt / Daubechies 1988): h = [1+sqrt3, 3+sqrt3, 3-sqrt3, 1-sqrt3] / (4 sqrt2)
_H = np.array([1 + np.sqrt(3), 3 + np.sqrt(3), 3 - np.sqrt(3), 1 - np.sqrt(3)]) / (4 * np.sqrt(2))
DB2_LO_D = _H[::-1].astype(np.float32)
DB2_HI_D = np.array([-_H[0], _H[1], -_H[2], _H[3]], np.float32)
DB2_LO_R = _H.astype(np.float32)
DB2_HI_R = DB2_HI_D[::-1].copy()


def sub_band_size(n, filter_len):
    """kernels.cu:150-152."""
    return (int(n) + int(filter_len) - 1) // 2


def _shift(s):
    return (ctypes.c_int * 3)(*[int(v) for v in (s if s is not None else (0, 0, 0))])


def _f(a):
    return np.ascontiguousarray(np.asarray(a, np.float32))


def cu_dwt3_axis(vol_data, decomp, vol_size, decomp_size, axis, lo_d, hi_d, rand_shift=(0, 0, 0), ctx=None):
    """cudwt3.rs:250-298: decomp <- one-level analysis of vol_data along axis."""
    V = _lib.as_cf32(vol_data, "vol_data")
    D = _lib.as_cf32(decomp, "decomp")
    lo, hi = _f(lo_d), _f(hi_d)
    if lo.size != hi.size:
        raise ValueError("filter lengths differ")
    _lib.check(_lib.load().llr_dwt3_axis(_lib.ctx_handle(ctx), _lib.ptr(V), _sz(V.size), _lib.ptr(D), _sz(D.size),
                                         _lib.sz3(vol_size), _lib.sz3(decomp_size), int(axis), _lib.ptr(lo),
                                         _lib.ptr(hi), _sz(lo.size), _shift(rand_shift)))


def cu_idwt3_axis(vol_data, decomp, vol_size, decomp_size, axis, lo_r, hi_r, rand_shift=(0, 0, 0), ctx=None):
    """cudwt3.rs:300-351: vol_data <- one-level synthesis of decomp along axis."""
    V = _lib.as_cf32(vol_data, "vol_data")
    D = _lib.as_cf32(decomp, "decomp")
    lo, hi = _f(lo_r), _f(hi_r)
    if lo.size != hi.size:
        raise ValueError("filter lengths differ")
    _lib.check(_lib.load().llr_idwt3_axis(_lib.ctx_handle(ctx), _lib.ptr(V), _sz(V.size), _lib.ptr(D), _sz(D.size),
                                          _lib.sz3(vol_size), _lib.sz3(decomp_size), int(axis), _lib.ptr(lo),
                                          _lib.ptr(hi), _sz(lo.size), _shift(rand_shift)))


def cu_dwt_denoise(vol_data, vol_size, lo_d, hi_d, lo_r, hi_r, rand_shift=(0, 0, 0), ctx=None):
    """cudwt3.rs:102-247: analysis along x, y, z then synthesis, in place."""
    V = _lib.as_cf32(vol_data, "vol_data")
    f = [_f(a) for a in (lo_d, hi_d, lo_r, hi_r)]
    if len({a.size for a in f}) != 1:
        raise ValueError("filter lengths differ")
    if V.size != int(np.prod(vol_size)):
        raise ValueError("volume length does not match its size")
    _lib.check(_lib.load().llr_dwt_denoise(_lib.ctx_handle(ctx), _lib.ptr(V), _lib.sz3(vol_size), *[_lib.ptr(a) for a in f],
                                           _sz(f[0].size), _shift(rand_shift)))
