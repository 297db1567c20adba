"""The allocation-free device signal model (llr_signal_model_device_noalloc)
against the allocating one (llr_signal_model_device), both through the C ABI
on the same device buffers: two calls in a row, the second on a smaller data
set, both directions, nonzero corrections. The tables are evaluated in FP64 on
the device instead of long double on the host, so single table entries may
differ in the last bit of FP32: bar 1e-6 relative."""
import numpy as np
import pytest

from llr_hip import _lib

pytestmark = pytest.mark.gpu

CASES = [((64, 48, 32), 6), ((32, 24, 16), 3)]


def _data(vol, nv, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(vol)) * nv
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    kc = rng.integers(-5, 6, (nv, 3)).astype(np.float32)
    ic = rng.uniform(-2.0, 2.0, (nv, 3)).astype(np.float32)
    return x, kc, ic


def test_noalloc_signal_model_matches_allocating_one(llr):
    import torch

    from llr_hip import signal_model as sm

    for it, (vol, nv) in enumerate(CASES):
        x, kc, ic = _data(vol, nv, 40 + it)
        for direction in (sm.ModelDirection.Forward, sm.ModelDirection.Inverse):
            a = torch.from_numpy(x).cuda()
            b = a.clone()
            torch.cuda.synchronize()
            sm.signal_model_device(a.data_ptr(), a.numel(), vol, nv, kc, ic, direction)
            sm.signal_model_device_noalloc(b.data_ptr(), b.numel(), vol, nv, kc, ic, direction)
            ra, rb = a.cpu().numpy(), b.cpu().numpy()
            e = float(np.linalg.norm((rb - ra).astype(np.complex128)) / np.linalg.norm(ra.astype(np.complex128)))
            print(f"[noalloc call {it + 1}: {vol} x{nv} {direction}] rel diff {e:.2e}", flush=True)
            assert np.all(np.isfinite(rb))
            assert e <= 1e-6


def test_noalloc_rejects_tables_beyond_the_scratch(llr):
    """400 volumes of 512x512x512 need 400 * (2 * 1536 + 3) + 1536 > 262144 table
    entries: rejected on the host before any device work (no data is touched)."""
    from llr_hip import signal_model as sm

    vol, nv = (512, 512, 512), 400
    with pytest.raises(_lib.LlrError) as ei:
        z = np.zeros((nv, 3), np.float32)
        sm.signal_model_device_noalloc(0x1000, int(np.prod(vol)) * nv, vol, nv, z, z, sm.ModelDirection.Forward)
    assert ei.value.status == _lib.LLR_ERR_UNSUPPORTED
    assert "scratch array holds 262144" in str(ei.value)
