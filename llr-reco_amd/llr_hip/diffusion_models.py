"""Diffusion-tensor scalars (diffusion_models.rs:21-253) on the GPU.

calc_scalars fits the log-linear tensor per voxel and returns its eigenvalues
and FA / MD / RD, computed by the HIP library (FP64 fit and eigen-solve).
"""
import ctypes

import numpy as np

from . import _lib

SCALAR_NAMES = ("e1", "e2", "e3", "fa", "md", "rd")


def _tables(bvals, bvecs, nq):
    b = np.ascontiguousarray(np.asarray(bvals, np.float32).ravel())
    g = np.ascontiguousarray(np.asarray(bvecs, np.float32).reshape(-1, 3))
    if b.size != nq or g.shape[0] != nq:
        raise ValueError("b-table length must equal the number of volumes")
    return b, g


def calc_scalars(volumes, n_voxels, bvals, bvecs, b0_tol=1e-6, ctx=None):
    """volumes: flat complex64 (n_volumes * n_voxels, volume slowest).
    Returns a dict of float32 arrays (n_voxels,) keyed e1, e2, e3, fa, md, rd."""
    v = _lib.as_cf32(volumes, "volumes")
    n_voxels = int(n_voxels)
    if v.size % n_voxels:
        raise ValueError("volumes length must be a multiple of n_voxels")
    nq = v.size // n_voxels
    b, g = _tables(bvals, bvecs, nq)
    out = np.empty(6 * n_voxels, np.float32)
    _lib.check(_lib.load().llr_dti_scalars(_lib.ctx_handle(ctx), _lib.ptr(v), ctypes.c_size_t(n_voxels),
                                           ctypes.c_size_t(nq), _lib.ptr(b), _lib.ptr(g), ctypes.c_float(b0_tol),
                                           _lib.ptr(out)))
    o = out.reshape(6, n_voxels)
    return {k: o[i] for i, k in enumerate(SCALAR_NAMES)}


def calc_scalars_device(d_volumes, n_voxels, n_volumes, bvals, bvecs, d_out, b0_tol=1e-6, ctx=None):
    """Device-pointer form: d_volumes (n_volumes * n_voxels complex64) and d_out
    (6 * n_voxels float32) are device addresses (ints)."""
    b, g = _tables(bvals, bvecs, int(n_volumes))
    _lib.check(_lib.load().llr_dti_scalars_device(
        _lib.ctx_handle(ctx), ctypes.c_void_p(int(d_volumes)), ctypes.c_size_t(int(n_voxels)),
        ctypes.c_size_t(int(n_volumes)), _lib.ptr(b), _lib.ptr(g), ctypes.c_float(b0_tol),
        ctypes.c_void_p(int(d_out))))
