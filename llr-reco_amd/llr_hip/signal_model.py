"""signal_model.rs / cufft.rs API on the GPU (signal_model.rs:14-117, cufft.rs:124-180)."""
import ctypes
import enum

import numpy as np

from . import _lib

_sz = ctypes.c_size_t


class ModelDirection(enum.IntEnum):
    """signal_model.rs:14-17."""
    Forward = _lib.LLR_FORWARD
    Inverse = _lib.LLR_INVERSE


class NormalizationType(enum.IntEnum):
    """cufft.rs:19-23."""
    NONE = _lib.LLR_NORM_NONE
    Inverse = _lib.LLR_NORM_INVERSE
    Unitary = _lib.LLR_NORM_UNITARY


def _corr(c, batch):
    a = np.ascontiguousarray(np.asarray(c, np.float32).reshape(-1))
    if a.size != 3 * batch:
        raise ValueError("the number of linear corrections must be equal to batch size")
    return a


def signal_model(x, image_size, batch_size, lin_ksp_corr, lin_img_corr, direction, ctx=None):
    """signal_model.rs:34-79 on a host array of batch_size volumes (in place)."""
    X = _lib.as_cf32(x, "x")
    kc = _corr(lin_ksp_corr, batch_size)
    ic = _corr(lin_img_corr, batch_size)
    _lib.check(_lib.load().llr_signal_model(_lib.ctx_handle(ctx), _lib.ptr(X), _sz(X.size), _lib.sz3(image_size),
                                            _sz(batch_size), _lib.ptr(kc), _lib.ptr(ic), int(direction)))


def signal_model_device(x_dev_ptr, n_elems, image_size, batch_size, lin_ksp_corr, lin_img_corr, direction, ctx=None):
    """signal_model on device memory (x_dev_ptr: integer device address)."""
    kc = _corr(lin_ksp_corr, batch_size)
    ic = _corr(lin_img_corr, batch_size)
    _lib.check(_lib.load().llr_signal_model_device(_lib.ctx_handle(ctx), ctypes.c_void_p(int(x_dev_ptr)),
                                                   _sz(n_elems), _lib.sz3(image_size), _sz(batch_size), _lib.ptr(kc),
                                                   _lib.ptr(ic), int(direction)))


def signal_model_device_noalloc(x_dev_ptr, n_elems, image_size, batch_size, lin_ksp_corr, lin_img_corr, direction,
                                ctx=None):
    """signal_model_device without any device allocation: the twiddle and ramp
    tables are built in the library's fixed 2 MiB per-GPU array (calls that
    overlap on one GPU race on it)."""
    kc = _corr(lin_ksp_corr, batch_size)
    ic = _corr(lin_img_corr, batch_size)
    _lib.check(_lib.load().llr_signal_model_device_noalloc(_lib.ctx_handle(ctx), ctypes.c_void_p(int(x_dev_ptr)),
                                                           _sz(n_elems), _lib.sz3(image_size), _sz(batch_size),
                                                           _lib.ptr(kc), _lib.ptr(ic), int(direction)))


def signal_model_batched(x, image_size, nq, batch_size, lin_ksp_corr, lin_img_corr, model_direction, ctx=None):
    """signal_model.rs:83-94: chunks of batch_size volumes. Each chunk gets its
    own slice of both correction lists (the reference passes the whole image
    correction list to every chunk, signal_model.rs:89-92; identical when the
    image corrections are all zero, which is the only case lib.rs uses)."""
    X = _lib.as_cf32(x, "x")
    N = int(np.prod(image_size))
    kc = np.asarray(lin_ksp_corr, np.float32).reshape(nq, 3)
    ic = np.asarray(lin_img_corr, np.float32).reshape(nq, 3)
    flat = X.reshape(-1, order="A")
    for s in range(0, nq, batch_size):
        e = min(nq, s + batch_size)
        chunk = flat[s * N:e * N]
        signal_model(chunk, image_size, e - s, kc[s:e], ic[s:e], model_direction, ctx)


def estimate_linear_correction(ksp, vol_size, ctx=None):
    """signal_model.rs:100-106 (arg-max |k|^2, lowest index on ties)."""
    K = _lib.as_cf32(ksp, "ksp")
    out = np.zeros(3, np.float32)
    _lib.check(_lib.load().llr_estimate_linear_correction(_lib.ctx_handle(ctx), _lib.ptr(K), _sz(K.size),
                                                          _lib.sz3(vol_size), _lib.ptr(out)))
    return out


def cu_fftn_batch(dims, batches, fft_direction, norm, data, ctx=None):
    """cufft.rs:124-180: batched 3-D C2C of col-major volumes, in place."""
    X = _lib.as_cf32(data, "data")
    _lib.check(_lib.load().llr_fftn_batch(_lib.ctx_handle(ctx), _lib.sz3(dims), _sz(batches), int(fft_direction),
                                          int(norm), _lib.ptr(X), _sz(X.size)))


def _fft3c(x, dims, batch_size, direction, vols_per_batch, ctx):
    X = _lib.as_cf32(x, "x")
    if X.size != int(np.prod(dims)) * batch_size:
        raise ValueError("dims must agree with the size of x")
    _lib.check(_lib.load().llr_fft3c(_lib.ctx_handle(ctx), _lib.ptr(X), _sz(X.size), _lib.sz3(dims), _sz(batch_size),
                                     int(direction), _sz(vols_per_batch)))


def fft3c(x, dims, batch_size=1, ctx=None):
    """cufft.rs:340-346: unitary forward centered FFT of each volume, in place."""
    _fft3c(x, dims, batch_size, _lib.LLR_FORWARD, 0, ctx)


def ifft3c(x, dims, batch_size=1, ctx=None):
    """cufft.rs:348-353: unitary inverse centered FFT of each volume, in place."""
    _fft3c(x, dims, batch_size, _lib.LLR_INVERSE, 0, ctx)


def fft3c_chunked(x, dims, batch_size, vols_per_batch, ctx=None):
    """cufft.rs:385-387 (vols_per_batch volumes through the device at a time)."""
    _fft3c(x, dims, batch_size, _lib.LLR_FORWARD, vols_per_batch, ctx)


def ifft3c_chunked(x, dims, batch_size, vols_per_batch, ctx=None):
    """cufft.rs:389-391."""
    _fft3c(x, dims, batch_size, _lib.LLR_INVERSE, vols_per_batch, ctx)


def cu_fft3(x, dims, direction, ctx=None):
    """cufft.rs:213-267: one volume, unnormalised forward, 1/N on the inverse."""
    X = _lib.as_cf32(x, "x")
    _lib.check(_lib.load().llr_cu_fft3(_lib.ctx_handle(ctx), _lib.ptr(X), _sz(X.size), _lib.sz3(dims), int(direction)))


def read_affine_translation(files, multiplier=1.0):
    """signal_model.rs:149-176: the translation (last 3 of the 12 'Parameters:'
    entries) of each affine text file, times multiplier; (n, 3) float32."""
    names = [str(f).encode() for f in files]
    arr = (ctypes.c_char_p * len(names))(*names)
    out = np.zeros((len(names), 3), np.float32)
    _lib.check(_lib.load().llr_read_affine_translation(arr, _sz(len(names)), ctypes.c_float(multiplier),
                                                       _lib.ptr(out)))
    return out


def read_affine_translation_range(n_range, file_pattern, multiplier=1.0):
    """signal_model.rs:149-161: file_pattern(i) for i in n_range."""
    return read_affine_translation([file_pattern(i) for i in n_range], multiplier)
