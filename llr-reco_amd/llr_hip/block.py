"""block.rs API on the GPU: shifted-block gather/scatter of Casorati matrices.

Mirrors lift_data_set / un_lift_data_set / grid_dim (block.rs:112-184). The
Rust functions take a slice of per-volume slices; here `volumes` is one
contiguous array holding every volume back to back (the layout lib.rs:374-388
slices), any fixed-size element type.
"""
import ctypes

import numpy as np

from . import _lib


def grid_dim(vol_size, block_size):
    """block.rs:112-118: number of whole blocks along each axis."""
    g = (ctypes.c_size_t * 3)()
    _lib.load().llr_grid_dim(_lib.sz3(vol_size), _lib.sz3(block_size), g)
    return [g[0], g[1], g[2]]


def _check(volumes, matrix_data):
    if not isinstance(volumes, np.ndarray) or not isinstance(matrix_data, np.ndarray):
        raise TypeError("volumes and matrix_data must be numpy arrays")
    if volumes.dtype != matrix_data.dtype:
        raise TypeError("volumes and matrix_data must share the element type")
    for a in (volumes, matrix_data):
        if not (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]):
            raise ValueError("arrays must be contiguous")


def lift_data_set(volumes, n_volumes, matrix_data, vol_size, block_size, shift, matrix_idx_start, n_matrices,
                  ctx=None):
    """block.rs:147-164: gather n_matrices blocks into (prod(block) x n_volumes) matrices."""
    _check(volumes, matrix_data)
    if volumes.size != int(np.prod(vol_size)) * n_volumes:
        raise ValueError("volumes must hold n_volumes * prod(vol_size) elements")
    _lib.check(_lib.load().llr_lift_data_set(
        _lib.ctx_handle(ctx), _lib.ptr(volumes), ctypes.c_size_t(n_volumes), ctypes.c_size_t(volumes.itemsize),
        _lib.ptr(matrix_data), ctypes.c_size_t(matrix_data.size), _lib.sz3(vol_size), _lib.sz3(block_size),
        _lib.sz3(shift), ctypes.c_size_t(matrix_idx_start), ctypes.c_size_t(n_matrices)))


def un_lift_data_set(volumes, n_volumes, matrix_data, vol_size, block_size, shift, matrix_idx_start, n_matrices,
                     ctx=None):
    """block.rs:166-184: scatter the matrices back into the volumes (in place)."""
    _check(volumes, matrix_data)
    if volumes.size != int(np.prod(vol_size)) * n_volumes:
        raise ValueError("volumes must hold n_volumes * prod(vol_size) elements")
    _lib.check(_lib.load().llr_un_lift_data_set(
        _lib.ctx_handle(ctx), _lib.ptr(volumes), ctypes.c_size_t(n_volumes), ctypes.c_size_t(volumes.itemsize),
        _lib.ptr(matrix_data), ctypes.c_size_t(matrix_data.size), _lib.sz3(vol_size), _lib.sz3(block_size),
        _lib.sz3(shift), ctypes.c_size_t(matrix_idx_start), ctypes.c_size_t(n_matrices)))


def lift_data_set_device(volumes_ptr, n_volumes, elem_bytes, matrix_ptr, matrix_len, vol_size, block_size, shift,
                         matrix_idx_start, n_matrices, ctx=None):
    """lift_data_set on device buffers (pointers on the context's device; matrix_len in elements)."""
    _lib.check(_lib.load().llr_lift_data_set_device(
        _lib.ctx_handle(ctx), ctypes.c_void_p(int(volumes_ptr)), ctypes.c_size_t(n_volumes),
        ctypes.c_size_t(elem_bytes), ctypes.c_void_p(int(matrix_ptr)), ctypes.c_size_t(matrix_len),
        _lib.sz3(vol_size), _lib.sz3(block_size), _lib.sz3(shift), ctypes.c_size_t(matrix_idx_start),
        ctypes.c_size_t(n_matrices)))


def un_lift_data_set_device(volumes_ptr, n_volumes, elem_bytes, matrix_ptr, matrix_len, vol_size, block_size, shift,
                            matrix_idx_start, n_matrices, ctx=None):
    """un_lift_data_set on device buffers (pointers on the context's device; matrix_len in elements)."""
    _lib.check(_lib.load().llr_un_lift_data_set_device(
        _lib.ctx_handle(ctx), ctypes.c_void_p(int(volumes_ptr)), ctypes.c_size_t(n_volumes),
        ctypes.c_size_t(elem_bytes), ctypes.c_void_p(int(matrix_ptr)), ctypes.c_size_t(matrix_len),
        _lib.sz3(vol_size), _lib.sz3(block_size), _lib.sz3(shift), ctypes.c_size_t(matrix_idx_start),
        ctypes.c_size_t(n_matrices)))
