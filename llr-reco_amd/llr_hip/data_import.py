"""Data ingest and egress of llr_recon_exec through the C ABI (host-only calls):
BART .cfl/.hdr volumes, the DTI data-set builder (data_import.rs:13-273), the
b-table and b0 mask, LlrReconParams TOML files (lib.rs:47-80) and
error_history.json (lib.rs:345-351). Arrays are numpy, column-major in the
reference's sense (x fastest); complex data is complex64."""
import ctypes

import numpy as np

from . import _lib

_sz = ctypes.c_size_t

DTI, B0, ALL = _lib.LLR_VOLUMES_DTI, _lib.LLR_VOLUMES_B0, _lib.LLR_VOLUMES_ALL


def _b(path):
    return str(path).encode()


def cfl_dims(path):
    """The sizes a .hdr holds (cfl::get_dims)."""
    d = (_sz * 16)()
    n = _sz()
    _lib.check(_lib.load().llr_cfl_dims(_b(path), d, ctypes.byref(n)))
    return [int(d[i]) for i in range(n.value)]


def read_cfl(path):
    """(flat complex64 data, dims) of a .cfl/.hdr pair (cfl::read_to_buffer)."""
    dims = cfl_dims(path)
    out = np.empty(int(np.prod(dims)), np.complex64)
    _lib.check(_lib.load().llr_cfl_read(_b(path), _lib.ptr(out), _sz(out.size)))
    return out, dims


def write_cfl(path, data, dims):
    """cfl::write_buffer: data flat, column-major, complex64."""
    a = np.ascontiguousarray(np.asarray(data, np.complex64).reshape(-1))
    d = (_sz * len(dims))(*[int(v) for v in dims])
    _lib.check(_lib.load().llr_cfl_write(_b(path), d, _sz(len(dims)), _lib.ptr(a), _sz(a.size)))


def read_b_table(path):
    """(b_vals (n,), b_vecs (n, 3)) of a b-table .txt (data_import.rs:110-145)."""
    n = _sz()
    L = _lib.load()
    _lib.check(L.llr_read_b_table(_b(path), None, None, _sz(0), ctypes.byref(n)))
    bv = np.zeros(n.value, np.float32)
    g = np.zeros((n.value, 3), np.float32)
    _lib.check(L.llr_read_b_table(_b(path), _lib.ptr(bv), _lib.ptr(g), _sz(n.value), ctypes.byref(n)))
    return bv, g


def b0_mask(b_vecs, tolerance):
    """data_import.rs:322-328: |g| <= tolerance per row."""
    g = np.ascontiguousarray(np.asarray(b_vecs, np.float32).reshape(-1, 3))
    m = np.zeros(g.shape[0], np.int32)
    _lib.check(_lib.load().llr_b0_mask(_lib.ptr(g), _sz(g.shape[0]), ctypes.c_float(tolerance), _lib.ptr(m)))
    return m.astype(bool)


class DtiDataSetBuilder:
    """data_import.rs:13-273 over a directory of per-volume .cfl files."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def from_cfl_volume_dir(cls, cfl_dir, filename_pattern, num_expected_images=None):
        h = ctypes.c_void_p()
        _lib.check(_lib.load().llr_dataset_from_cfl_dir(_b(cfl_dir), _b(filename_pattern),
                                                        _sz(num_expected_images or 0), ctypes.byref(h)))
        return cls(h)

    def with_b_table(self, b_table, b0_tol=0.0):
        _lib.check(_lib.load().llr_dataset_with_b_table(self.handle, _b(b_table), ctypes.c_float(b0_tol)))
        return self

    def with_sample_mask(self, sample_mask_cfl):
        _lib.check(_lib.load().llr_dataset_with_sample_mask(self.handle, _b(sample_mask_cfl)))
        return self

    def info(self):
        v = (_sz * 3)()
        n, nd, nb = _sz(), _sz(), _sz()
        _lib.check(_lib.load().llr_dataset_info(self.handle, v, ctypes.byref(n), ctypes.byref(nd), ctypes.byref(nb)))
        return dict(vol_dims=tuple(v), total_volumes=n.value, n_dti=nd.value, n_b0=nb.value)

    def total_volumes(self):
        return self.info()["total_volumes"]

    def file(self, i):
        buf = ctypes.create_string_buffer(4096)
        _lib.check(_lib.load().llr_dataset_file(self.handle, _sz(i), buf, _sz(4096)))
        return buf.value.decode()

    def _indices(self, which):
        n = _sz()
        L = _lib.load()
        _lib.check(L.llr_dataset_indices(self.handle, int(which), None, _sz(0), ctypes.byref(n)))
        out = np.zeros(n.value, np.uint64)
        _lib.check(L.llr_dataset_indices(self.handle, int(which), _lib.ptr(out), _sz(n.value), ctypes.byref(n)))
        return [int(i) for i in out]

    def dti_indices(self):
        return self._indices(DTI)

    def b0_indices(self):
        return self._indices(B0)

    def b_table(self):
        n = self.total_volumes()
        bv = np.zeros(n, np.float32)
        g = np.zeros((n, 3), np.float32)
        _lib.check(_lib.load().llr_dataset_b_table(self.handle, _lib.ptr(bv), _lib.ptr(g)))
        return bv, g

    def _load(self, which):
        vol = self.info()["vol_dims"]
        n = len(self._indices(which)) if which != ALL else self.total_volumes()
        out = np.empty(int(np.prod(vol)) * n, np.complex64)
        _lib.check(_lib.load().llr_dataset_load(self.handle, int(which), _lib.ptr(out), _sz(out.size)))
        return out

    def load_dti_set(self):
        return self._load(DTI)

    def load_b0_set(self):
        return self._load(B0)

    def load_all(self):
        return self._load(ALL)

    def _mask(self, which):
        vol = self.info()["vol_dims"]
        n = len(self._indices(which)) if which != ALL else self.total_volumes()
        out = np.empty(vol[1] * vol[2] * n, np.complex64)
        _lib.check(_lib.load().llr_dataset_sample_mask(self.handle, int(which), _lib.ptr(out), _sz(out.size)))
        return out

    def dti_sample_mask(self):
        return self._mask(DTI)

    def b0_sample_mask(self):
        return self._mask(B0)

    def full_sample_mask(self):
        return self._mask(ALL)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().llr_dataset_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PARAM_FIELDS = ("llr_block_size", "max_rand_shift", "fft_max_batch_size", "cas_mat_max_batch_size", "llr_block_rank",
                "wavelet_shrink_factor", "n_iter", "subspace_projection_mat")


def _params_to_dict(p):
    d = dict(llr_block_size=tuple(p.llr_block_size), max_rand_shift=tuple(p.max_rand_shift) if p.has_max_rand_shift
             else None, fft_max_batch_size=p.fft_max_batch_size, cas_mat_max_batch_size=p.cas_mat_max_batch_size,
             llr_block_rank=p.llr_block_rank, wavelet_shrink_factor=p.wavelet_shrink_factor, n_iter=p.n_iter,
             subspace_projection_mat=p.subspace_projection_mat.decode() if p.has_subspace else None,
             llr_mode={_lib.LLR_MODE_RANK: "rank", _lib.LLR_MODE_SVT: "svt"}[p.llr_mode],
             svt_lambda_factor=p.lambda_factor, precision={0: "fp32", 1: "mixed"}[p.precision],
             storage={0: "fp32", 1: "fp16"}[p.storage], shift_seed=p.seed)
    return d


def default_recon_params():
    """LlrReconParams::default (lib.rs:67-80) plus this library's added keys."""
    p = _lib.TomlParams()
    _lib.load().llr_recon_params_default(ctypes.byref(p))
    return _params_to_dict(p)


def read_recon_params(path):
    """<path>.toml as bin/llr-recon.rs:8-12 reads it."""
    p = _lib.TomlParams()
    _lib.check(_lib.load().llr_recon_params_read(_b(path), ctypes.byref(p)))
    return _params_to_dict(p)


def write_recon_params(path, params=None):
    """<path>.toml with the reference's fields (bin/llr-reco-params.rs:14-23)."""
    p = _lib.TomlParams()
    L = _lib.load()
    L.llr_recon_params_default(ctypes.byref(p))
    if params:
        d = dict(params)
        p.llr_block_size[:] = list(d.get("llr_block_size", p.llr_block_size))
        if d.get("max_rand_shift") is not None:
            p.has_max_rand_shift = 1
            p.max_rand_shift[:] = list(d["max_rand_shift"])
        for k in ("fft_max_batch_size", "cas_mat_max_batch_size", "llr_block_rank", "n_iter"):
            if k in d:
                setattr(p, k, int(d[k]))
        if "wavelet_shrink_factor" in d:
            p.wavelet_shrink_factor = float(d["wavelet_shrink_factor"])
        if "subspace_projection_mat" in d:
            p.has_subspace = 1 if d["subspace_projection_mat"] else 0
            p.subspace_projection_mat = (d["subspace_projection_mat"] or "").encode()
    _lib.check(L.llr_recon_params_write(_b(path), ctypes.byref(p)))


def write_error_history(path, hist):
    h = np.ascontiguousarray(np.asarray(hist, np.float64))
    _lib.check(_lib.load().llr_write_error_history(_b(path), _lib.ptr(h), _sz(h.size)))
