"""low_rank.rs / cublas.rs API on the GPU (low_rank.rs:10-69, cublas.rs:12-127)."""
import ctypes

import numpy as np

from . import _lib

_sz = ctypes.c_size_t


def cu_low_rank_approx_batch(m, n, rank, batch_size, matrix_data, ctx=None):
    """low_rank.rs:10-69: rank-`rank` truncation of each m x n matrix, in place."""
    A = _lib.as_cf32(matrix_data, "matrix_data")
    _lib.check(_lib.load().llr_low_rank_approx_batch(_lib.ctx_handle(ctx), _sz(m), _sz(n), _sz(rank),
                                                     _sz(batch_size), _lib.ptr(A), _sz(A.size)))


def svt_batch(m, n, lam, batch_size, matrix_data, ctx=None):
    """Singular-value soft threshold of each m x n matrix, in place. Returns sigma_max."""
    A = _lib.as_cf32(matrix_data, "matrix_data")
    smax = ctypes.c_float()
    _lib.check(_lib.load().llr_svt_batch(_lib.ctx_handle(ctx), m, n, float(lam), batch_size, _lib.ptr(A), A.size,
                                         ctypes.byref(smax)))
    return smax.value


def cublas_cgemm_strided_batched(m, k, n, batch_size, broadcast_a, alpha, beta, transpose_a, transpose_b, a, b, c,
                                 ctx=None):
    """cublas.rs:12-105 on host arrays: c = alpha*op(a)*op(b) + beta*c (in place in c)."""
    A = _lib.as_cf32(a, "a")
    B = _lib.as_cf32(b, "b")
    Cm = _lib.as_cf32(c, "c")
    al = _lib.CF32(float(np.real(alpha)), float(np.imag(alpha)))
    be = _lib.CF32(float(np.real(beta)), float(np.imag(beta)))
    _lib.check(_lib.load().llr_cgemm_strided_batched(_lib.ctx_handle(ctx), m, k, n, batch_size, int(broadcast_a), al,
                                                     be, int(transpose_a), int(transpose_b), _lib.ptr(A), A.size,
                                                     _lib.ptr(B), B.size, _lib.ptr(Cm), Cm.size))


def rhs_cdgmm_batched(m, n, batch_size, matrix, diag, ctx=None):
    """cublas.rs:119-127 / kernels.cu:11-67: matrix[:, j, b] *= diag[j, b], in place."""
    A = _lib.as_cf32(matrix, "matrix")
    d = np.ascontiguousarray(diag, np.float32)
    if d.size < n * batch_size:
        raise ValueError("diag needs n * batch_size entries")
    _lib.check(_lib.load().llr_rhs_cdgmm_batched(_lib.ctx_handle(ctx), _sz(m), _sz(n), _sz(batch_size), _lib.ptr(A),
                                                 _sz(A.size), _lib.ptr(d)))
