"""dti_subspace.rs API on the GPU (dti_subspace.rs:89-221).

Matrices are complex64, column-major, batched back to back: image matrices are
nvox x nq, compressed matrices srank x nvox, the subspace nq x srank.
"""
import ctypes

from . import _lib

_sz = ctypes.c_size_t


def dti_project(subspace_matrix, image_matrices, compressed_matrices, batch_size, nq, nvox, srank, ctx=None):
    """dti_subspace.rs:141-166: compressed = S^H X^H per matrix."""
    S = _lib.as_cf32(subspace_matrix, "subspace_matrix")
    X = _lib.as_cf32(image_matrices, "image_matrices")
    Cm = _lib.as_cf32(compressed_matrices, "compressed_matrices")
    _lib.check(_lib.load().llr_dti_project(_lib.ctx_handle(ctx), _lib.ptr(S), _sz(S.size), _lib.ptr(X), _sz(X.size),
                                           _lib.ptr(Cm), _sz(Cm.size), _sz(batch_size), _sz(nq), _sz(nvox),
                                           _sz(srank)))


def dti_back_project(subspace_matrix, image_matrices, compressed_matrices, batch_size, nq, nvox, srank, ctx=None):
    """dti_subspace.rs:172-207: image = (S * compressed)^H per matrix (in place)."""
    S = _lib.as_cf32(subspace_matrix, "subspace_matrix")
    X = _lib.as_cf32(image_matrices, "image_matrices")
    Cm = _lib.as_cf32(compressed_matrices, "compressed_matrices")
    _lib.check(_lib.load().llr_dti_back_project(_lib.ctx_handle(ctx), _lib.ptr(S), _sz(S.size), _lib.ptr(X),
                                                _sz(X.size), _lib.ptr(Cm), _sz(Cm.size), _sz(batch_size), _sz(nq),
                                                _sz(nvox), _sz(srank)))


def conj_transpose_matrix(matrix_data, m, n, batch_size, ctx=None):
    """dti_subspace.rs:211-221: (m x n) -> (n x m)^H per matrix, in place."""
    A = _lib.as_cf32(matrix_data, "matrix_data")
    _lib.check(_lib.load().llr_conj_transpose_matrix(_lib.ctx_handle(ctx), _lib.ptr(A), _sz(A.size), _sz(m), _sz(n),
                                                     _sz(batch_size)))


def subspace_constraint(subspace_data, image_matrices, batch_size, nq, nvox, srank, ctx=None):
    """dti_subspace.rs:89-134 (in place)."""
    S = _lib.as_cf32(subspace_data, "subspace_data")
    X = _lib.as_cf32(image_matrices, "image_matrices")
    _lib.check(_lib.load().llr_subspace_constraint(_lib.ctx_handle(ctx), _lib.ptr(S), _sz(S.size), _lib.ptr(X),
                                                   _sz(X.size), _sz(batch_size), _sz(nq), _sz(nvox), _sz(srank)))
