"""fftshift.rs API on the GPU (fftshift.rs:8-116)."""
import ctypes
import enum

import numpy as np

from . import _lib


class PhaseShiftDir(enum.IntEnum):
    """fftshift.rs:8-11."""
    Forward = _lib.LLR_FORWARD
    Inverse = _lib.LLR_INVERSE


def _dims3(dims):
    d = [int(x) for x in dims]
    if len(d) > 3:
        raise ValueError("greater than 3 dimensions is not supported here")
    return d + [1] * (3 - len(d))


def fftshift(dims, data, ctx=None):
    """fftshift.rs:103-108: circular shift by floor(d/2) per axis, in place."""
    if not isinstance(data, np.ndarray) or not (data.flags["C_CONTIGUOUS"] or data.flags["F_CONTIGUOUS"]):
        raise TypeError("data must be a contiguous numpy array")
    d = _dims3(dims)
    if int(np.prod(d)) != data.size:
        raise ValueError("dims and source size must agree")
    _lib.check(_lib.load().llr_fftshift(_lib.ctx_handle(ctx), _lib.ptr(data), ctypes.c_size_t(data.itemsize),
                                        _lib.sz3(d)))


def ifftshift(dims, data, ctx=None):
    """fftshift.rs:111-116: circular shift by ceil(d/2) per axis, in place."""
    if not isinstance(data, np.ndarray) or not (data.flags["C_CONTIGUOUS"] or data.flags["F_CONTIGUOUS"]):
        raise TypeError("data must be a contiguous numpy array")
    d = _dims3(dims)
    if int(np.prod(d)) != data.size:
        raise ValueError("dims and source size must agree")
    _lib.check(_lib.load().llr_ifftshift(_lib.ctx_handle(ctx), _lib.ptr(data), ctypes.c_size_t(data.itemsize),
                                         _lib.sz3(d)))


def phase_shift3(dims, x, direction, ctx=None):
    """fftshift.rs:42-74: FOV/2 phase modulation, in place."""
    X = _lib.as_cf32(x, "x")
    d = _dims3(dims)
    _lib.check(_lib.load().llr_phase_shift3(_lib.ctx_handle(ctx), _lib.ptr(X), ctypes.c_size_t(X.size), _lib.sz3(d),
                                            int(direction)))
