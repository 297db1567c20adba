"""Warm-start 3D U-Net (llr_unet_*): BF16 MFMA forward/backward kernels and
Adam on the GPU. Tensors are numpy (host) or device pointers, shaped
(nz, ny, nx, channels) with channels fastest."""
import ctypes
import sys

import numpy as np

from . import _lib

_sz = ctypes.c_size_t


def _dims(shape):
    nz, ny, nx = shape[:3]
    return (ctypes.c_size_t * 3)(int(nx), int(ny), int(nz))


class UNet:
    def __init__(self, in_channels=16, base_channels=16, levels=4, seed=0, ctx=None):
        self.ctx = ctx or _lib.default_context()
        self.cin, self.base, self.levels = int(in_channels), int(base_channels), int(levels)
        h = ctypes.c_void_p()
        _lib.check(_lib.load().llr_unet_create(self.ctx.handle, self.cin, self.base, self.levels,
                                               ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self.handle = h
        n = _sz()
        _lib.check(_lib.load().llr_unet_num_params(self.handle, ctypes.byref(n)))
        self.n_params = n.value

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        _lib.check(_lib.load().llr_unet_get_params(self.handle, _lib.ptr(p), _sz(p.size)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _lib.check(_lib.load().llr_unet_set_params(self.handle, _lib.ptr(p), _sz(p.size)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty_like(x)
        _lib.check(_lib.load().llr_unet_forward(self.handle, _lib.ptr(x), _lib.ptr(y), _dims(x.shape), 0))
        return y

    def forward_device(self, x_ptr, y_ptr, shape):
        _lib.check(_lib.load().llr_unet_forward(self.handle, ctypes.c_void_p(int(x_ptr)), ctypes.c_void_p(int(y_ptr)),
                                                _dims(shape), 1))

    def gradients(self, x, gy):
        x = np.ascontiguousarray(x, np.float32)
        gy = np.ascontiguousarray(gy, np.float32)
        g = np.empty(self.n_params, np.float32)
        _lib.check(_lib.load().llr_unet_gradients(self.handle, _lib.ptr(x), _lib.ptr(gy), _dims(x.shape), _lib.ptr(g)))
        return g

    def train_step(self, x, target, lr=1e-3):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = ctypes.c_double()
        _lib.check(_lib.load().llr_unet_train_step(self.handle, _lib.ptr(x), _lib.ptr(t), _dims(x.shape),
                                                   ctypes.c_float(lr), 0, ctypes.byref(loss)))
        return loss.value

    def train_step_device(self, x_ptr, t_ptr, shape, lr=1e-3):
        loss = ctypes.c_double()
        _lib.check(_lib.load().llr_unet_train_step(self.handle, ctypes.c_void_p(int(x_ptr)), ctypes.c_void_p(int(t_ptr)),
                                                   _dims(shape), ctypes.c_float(lr), 1, ctypes.byref(loss)))
        return loss.value

    ACT_KINDS = {"a": 0, "b": 1, "pool": 2, "cat": 3, "da": 4, "db": 5}

    def activation(self, kind, level, shape):
        """A stored activation of the last forward pass; shape (nz, ny, nx, c) at that level."""
        out = np.empty(int(np.prod(shape)), np.float32)
        _lib.check(_lib.load().llr_unet_get_activation(self.handle, self.ACT_KINDS[kind], int(level), _lib.ptr(out),
                                                       _sz(out.size)))
        return out.reshape(shape)

    def predict_coefficients_device(self, coef_ptr, K, vol, scale, out_ptr, to_net=None, from_net=None):
        """K complex planes (device) -> predicted planes (device; may alias)."""
        tn, fn = _mix(to_net, K), _mix(from_net, K)
        _lib.check(_lib.load().llr_unet_predict_coefficients(
            self.handle, ctypes.c_void_p(int(coef_ptr)), _sz(K), (ctypes.c_size_t * 3)(*[int(v) for v in vol]),
            ctypes.c_float(scale), _mptr(tn), _mptr(fn), ctypes.c_void_p(int(out_ptr))))

    def train_step_crop(self, x_ptr, t_ptr, full_shape, origin, crop_shape, lr=1e-3):
        """One Adam step on a crop of stored BF16 tensors (device, (nz, ny, nx, channels)).
        origin and crop_shape are (z, y, x)."""
        loss = ctypes.c_double()
        _lib.check(_lib.load().llr_unet_train_step_crop(
            self.handle, ctypes.c_void_p(int(x_ptr)), ctypes.c_void_p(int(t_ptr)), _dims(full_shape),
            (ctypes.c_size_t * 3)(int(origin[2]), int(origin[1]), int(origin[0])), _dims(crop_shape),
            ctypes.c_float(lr), ctypes.byref(loss)))
        return loss.value

    def save(self, path, **meta):
        """Parameters and the architecture (plus any metadata) as an .npz file."""
        np.savez_compressed(path, params=self.get_params(), cin=self.cin, base=self.base, levels=self.levels,
                            **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path, ctx=None):
        with np.load(path, allow_pickle=False) as f:
            net = cls(int(f["cin"]), int(f["base"]), int(f["levels"]), seed=0, ctx=ctx)
            net.set_params(f["params"])
            meta = {k: f[k] for k in f.files if k not in ("params", "cin", "base", "levels")}
        net.meta = meta
        return net

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().llr_unet_destroy(self.handle)
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


def _mix(m, K):
    if m is None:
        return None
    m = np.ascontiguousarray(np.asarray(m, np.float32))
    if m.shape != (K, K):
        raise ValueError(f"mixing matrix must be {K}x{K}")
    return m


def _mptr(m):
    return None if m is None else _lib.ptr(m)


def pack_coefficients_device(coef_ptr, K, vol, channels, scale, out_ptr, to_net=None, bf16=False, ctx=None):
    """K complex planes (device) -> (voxels, channels) network tensor (device), FP32 or BF16."""
    tn = _mix(to_net, K)
    _lib.check(_lib.load().llr_unet_pack_coefficients(
        _lib.ctx_handle(ctx), ctypes.c_void_p(int(coef_ptr)), _sz(K), (ctypes.c_size_t * 3)(*[int(v) for v in vol]),
        int(channels), ctypes.c_float(scale), _mptr(tn), int(bool(bf16)), ctypes.c_void_p(int(out_ptr))))


def coefficient_scale_device(coef_ptr, K, vol, to_net=None, ctx=None):
    """sqrt(mean over voxels of |(to_net c)_0|^2) of K complex planes (device)."""
    tn = _mix(to_net, K)
    out = ctypes.c_double()
    _lib.check(_lib.load().llr_unet_coefficient_scale(
        _lib.ctx_handle(ctx), ctypes.c_void_p(int(coef_ptr)), _sz(K), (ctypes.c_size_t * 3)(*[int(v) for v in vol]),
        _mptr(tn), ctypes.byref(out)))
    return out.value


def params_to_dict(p, cin=16, base=16, levels=4):
    """Split the flat parameter vector into {layer: [w (kk, co, ci), b (co)]} in the library's layer order."""
    from . import unet_layout
    out, off = {}, 0
    for name, k, ci, co, _ in unet_layout.layer_specs(cin, base, levels):
        nw = k ** 3 * co * ci
        out[name] = [p[off:off + nw].reshape(k ** 3, co, ci), p[off + nw:off + nw + co]]
        off += nw + co
    return out


def dict_to_params(P, cin=16, base=16, levels=4):
    from . import unet_layout
    parts = []
    for name, k, ci, co, _ in unet_layout.layer_specs(cin, base, levels):
        parts += [np.asarray(P[name][0], np.float32).ravel(), np.asarray(P[name][1], np.float32).ravel()]
    return np.concatenate(parts)
