"""ctypes loader for libllr_hip.so (the C ABI declared in include/llr_hip.h).

The library is required: importing any compute entry point without it raises,
there is no CPU fallback.
"""
import ctypes
import sys
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libllr_hip.so")

LLR_OK = 0
LLR_ERR_ARG = 1
LLR_ERR_HIP = 2
LLR_ERR_NOMEM = 3
LLR_ERR_UNSUPPORTED = 4
LLR_ERR_NUMERIC = 5
LLR_FORWARD = 0
LLR_INVERSE = 1
LLR_NORM_NONE = 0
LLR_NORM_INVERSE = 1
LLR_NORM_UNITARY = 2
LLR_MODE_RANK = 0
LLR_MODE_SVT = 1
LLR_PREC_FP32 = 0
LLR_PREC_MIXED = 1
LLR_STORE_FP32 = 0
LLR_STORE_FP16 = 1
ABI_VERSION = 108
LLR_NUM_STAGES = 16
# llr_recon_stage_times order; fft_x/fft_y are the forward passes, *_adjoint the
# inverse ones; wavelet_* are the wavelet's kernels (also inside "wavelet")
STAGE_NAMES = ["llr_block", "subspace_project", "wavelet", "fft_x", "fft_y", "fft_z_project", "reductions", "setup",
               "unet", "fft_y_adjoint", "fft_x_adjoint", "wavelet_level0_in", "wavelet_level0_mid",
               "wavelet_level0_out", "wavelet_deeper", "fft_y_adjoint_project"]

# every symbol include/llr_hip.h declares
EXPORTS = [
    "llr_last_error", "llr_abi_version", "llr_ctx_create", "llr_ctx_destroy", "llr_ctx_synchronize",
    "llr_ctx_device", "llr_grid_dim", "llr_lift_data_set", "llr_un_lift_data_set", "llr_lift_data_set_device",
    "llr_un_lift_data_set_device", "llr_dti_project",
    "llr_dti_back_project", "llr_conj_transpose_matrix", "llr_subspace_constraint", "llr_low_rank_approx_batch",
    "llr_svt_batch", "llr_cgemm_strided_batched", "llr_rhs_cdgmm_batched", "llr_signal_model",
    "llr_signal_model_device", "llr_estimate_linear_correction", "llr_fftn_batch", "llr_fftshift", "llr_ifftshift",
    "llr_phase_shift3", "llr_wavelet_denoise", "llr_hard_project", "llr_calc_undersampling_factor",
    "llr_under_sample_kspace", "llr_dti_scalars", "llr_dti_scalars_device", "llr_recon_create",
    "llr_recon_destroy", "llr_recon_set_measurements",
    "llr_recon_set_initial", "llr_recon_run", "llr_recon_get_image", "llr_recon_info", "llr_recon_mixed_stats",
    "llr_recon_hybrid",
    "llr_recon_enable_timing",
    "llr_recon_stage_times", "llr_recon_batch_create", "llr_recon_batch_destroy",
    "llr_recon_batch_set_measurements", "llr_recon_batch_run", "llr_recon_batch_get_image", "llr_recon_batch_tuning",
    "llr_unet_create", "llr_unet_destroy", "llr_unet_num_params", "llr_unet_get_params", "llr_unet_set_params",
    "llr_unet_forward", "llr_unet_gradients", "llr_unet_train_step", "llr_unet_predict_coefficients",
    "llr_unet_pack_coefficients", "llr_unet_get_activation", "llr_unet_coefficient_scale", "llr_unet_train_step_crop",
    "llr_recon_prepare", "llr_recon_get_coefficients", "llr_recon_set_coefficients", "llr_recon_warm_start",
    "llr_recon_run_learned", "llr_signal_model_device_noalloc",
    "llr_cfl_dims", "llr_cfl_read", "llr_cfl_write", "llr_read_b_table", "llr_b0_mask", "llr_dataset_from_cfl_dir",
    "llr_dataset_destroy", "llr_dataset_with_b_table", "llr_dataset_with_sample_mask", "llr_dataset_info",
    "llr_dataset_file", "llr_dataset_indices", "llr_dataset_b_table", "llr_dataset_load", "llr_dataset_sample_mask",
    "llr_recon_params_default", "llr_recon_params_read", "llr_recon_params_write", "llr_write_error_history",
    "llr_fft3c", "llr_cu_fft3", "llr_read_affine_translation", "llr_wavelet_denoise_shifted", "llr_dwt3_axis",
    "llr_idwt3_axis", "llr_dwt_denoise",
]

LLR_VOLUMES_DTI = 0
LLR_VOLUMES_B0 = 1
LLR_VOLUMES_ALL = 2


class TomlParams(ctypes.Structure):
    """llr_toml_params (include/llr_hip.h)."""
    _fields_ = [
        ("llr_block_size", ctypes.c_size_t * 3),
        ("has_max_rand_shift", ctypes.c_int),
        ("max_rand_shift", ctypes.c_size_t * 3),
        ("fft_max_batch_size", ctypes.c_size_t),
        ("cas_mat_max_batch_size", ctypes.c_size_t),
        ("llr_block_rank", ctypes.c_size_t),
        ("wavelet_shrink_factor", ctypes.c_float),
        ("n_iter", ctypes.c_size_t),
        ("has_subspace", ctypes.c_int),
        ("subspace_projection_mat", ctypes.c_char * 4096),
        ("llr_mode", ctypes.c_int),
        ("lambda_factor", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("storage", ctypes.c_int),
        ("seed", ctypes.c_ulonglong),
    ]


class LlrError(RuntimeError):
    """A nonzero llr_status from the library (the reference panics here)."""

    def __init__(self, status, msg):
        super().__init__(f"llr status {status}: {msg}")
        self.status = status


class CF32(ctypes.Structure):
    _fields_ = [("re", ctypes.c_float), ("im", ctypes.c_float)]


class WarmParams(ctypes.Structure):
    _fields_ = [
        ("to_net", ctypes.c_void_p),
        ("from_net", ctypes.c_void_p),
        ("residual_stop", ctypes.c_double),
        ("max_iter", ctypes.c_size_t),
        ("momentum", ctypes.c_double),
        ("min_iter", ctypes.c_size_t),
    ]


class ReconParams(ctypes.Structure):
    _fields_ = [
        ("vol", ctypes.c_size_t * 3),
        ("nq", ctypes.c_size_t),
        ("blk", ctypes.c_size_t * 3),
        ("n_iter", ctypes.c_size_t),
        ("llr_mode", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("rank", ctypes.c_size_t),
        ("svt_lambda", ctypes.c_double),
        ("lambda_factor", ctypes.c_double),
        ("wavelet_alpha", ctypes.c_double),
        ("wavelet_levels", ctypes.c_int),
        ("srank", ctypes.c_size_t),
        ("subspace", ctypes.c_void_p),
        ("shifts", ctypes.c_void_p),
        ("residual_stop", ctypes.c_double),
        ("partial_blocks", ctypes.c_int),
        ("storage", ctypes.c_int),
        ("no_hybrid", ctypes.c_int),
        ("lin_img_corr", ctypes.c_void_p),
        ("fused_projection", ctypes.c_int),
    ]


_lib = None
_lock = threading.Lock()


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                                  "(make -C llr-reco_amd/csrc); there is no CPU fallback")
            # torch ships its own copy of the HIP runtime under the same soname
            # (libamdhip64.so.7). Loaded first, it is the one this library binds
            # to, so a process that also uses torch for device memory has one
            # HIP runtime; loaded after, torch brings a second one, and this
            # library's runtime then failed to find the device once torch had
            # initialised it.
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            lib = ctypes.CDLL(LIB_PATH)
            lib.llr_last_error.restype = ctypes.c_char_p
            lib.llr_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
            lib.llr_ctx_destroy.argtypes = [ctypes.c_void_p]
            lib.llr_ctx_destroy.restype = None
            lib.llr_recon_destroy.argtypes = [ctypes.c_void_p]
            lib.llr_recon_destroy.restype = None
            lib.llr_grid_dim.restype = None
            lib.llr_dataset_destroy.argtypes = [ctypes.c_void_p]
            lib.llr_dataset_destroy.restype = None
            lib.llr_recon_params_default.restype = None
            lib.llr_svt_batch.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_float,
                                          ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
            lib.llr_wavelet_denoise.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                                ctypes.c_float, ctypes.c_int]
            lib.llr_cgemm_strided_batched.argtypes = [
                ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int,
                CF32, CF32, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
            _lib = lib
    return _lib


def last_error():
    return load().llr_last_error().decode(errors="replace")


def check(status):
    if status != 0:
        raise LlrError(status, last_error())


def sz3(v):
    v = [int(x) for x in v]
    if len(v) != 3:
        raise ValueError("expected 3 dimensions")
    return (ctypes.c_size_t * 3)(*v)


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def as_cf32(a, name="array"):
    a = np.asarray(a)
    if a.dtype != np.complex64:
        raise TypeError(f"{name} must be complex64 (Complex32), got {a.dtype}")
    if not (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]):
        raise ValueError(f"{name} must be contiguous")
    return a


class Context:
    """One HIP device + stream (llr_ctx)."""

    def __init__(self, device=0):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.llr_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        check(load().llr_ctx_synchronize(self.handle))

    def close(self):
        if self.handle:
            load().llr_ctx_destroy(self.handle)
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


_default_ctx = {}


def default_context(device=0):
    """Process-wide context per device, created on first use."""
    with _lock:
        c = _default_ctx.get(device)
    if c is None:
        c = Context(device)
        with _lock:
            _default_ctx[device] = c
    return c


def ctx_handle(ctx):
    return (ctx or default_context()).handle
