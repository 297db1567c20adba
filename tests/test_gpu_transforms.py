"""GPU entry points around the signal model and the wavelet that the
reconstruction loop does not call itself:
  - cu_dwt3_axis / cu_idwt3_axis / cu_dwt_denoise (cudwt3.rs:19-351,
    kernels.cu:155-349) with any filter pair and the circular shift hook,
    including the reference's perfect-reconstruction tests (cudwt3.rs:19-93)
    at reduced size;
  - the cycle-spinning multi-level shrinkage;
  - fft3c / ifft3c / the chunked forms / cu_fft3 (cufft.rs:213-402);
  - the reconstruction with nonzero image corrections (signal_model.rs:34-79).
"""
import numpy as np
import pytest

import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(123)


def crandn(n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)


def rel(a, b):
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(np.asarray(a, np.complex128) - b) / np.linalg.norm(b))


def ramp(vol):
    i, j, k = np.meshgrid(*[np.arange(n) for n in vol], indexing="ij")
    return (i + j + k).astype(np.float32).ravel(order="F").astype(np.complex64)


def roll3(x, vol, s):
    """np.roll of a flat column-major volume: out[i] = x[i - s] per axis."""
    v = x.reshape(vol[2], vol[1], vol[0])
    return np.roll(v, (s[2], s[1], s[0]), axis=(0, 1, 2)).ravel()


@pytest.mark.parametrize("vol", [(128, 128, 128), (96, 80, 50)])
def test_dwt3_axis_round_trip_kat(llr, vol):
    """cudwt3.rs:40-93 (512^3 there): a ramp i + j + k through the analysis and
    synthesis along each axis rounds back to itself."""
    from llr_hip import cudwt3 as W

    x0 = ramp(vol)
    for axis in range(3):
        dec = list(vol)
        dec[axis] = 2 * W.sub_band_size(vol[axis], 4)
        r = np.zeros(int(np.prod(dec)), np.complex64)
        W.cu_dwt3_axis(x0, r, vol, dec, axis, W.DB2_LO_D, W.DB2_HI_D)
        x = np.zeros_like(x0)
        W.cu_idwt3_axis(x, r, vol, dec, axis, W.DB2_LO_R, W.DB2_HI_R)
        err = np.abs(x - x0).max()
        print(f"[dwt3 axis {axis} {vol}] max error {err:.3e}", flush=True)
        assert err < 1e-3 * np.abs(x0).max()
        assert np.array_equal(np.round(x.real) + 1j * np.round(x.imag), x0)


def test_dwt_denoise_round_trip_kat(llr):
    """cudwt3.rs:19-37: the 3-axis analysis + synthesis (no shrinkage) is the
    identity, with and without a shift, on a non-cubic box."""
    from llr_hip import cudwt3 as W

    vol = (64, 48, 40)
    x0 = ramp(vol)
    for shift in ((0, 0, 0), (5, 7, 3), (-3, 20, 39)):
        x = x0.copy()
        W.cu_dwt_denoise(x, vol, W.DB2_LO_D, W.DB2_HI_D, W.DB2_LO_R, W.DB2_HI_R, shift)
        assert np.array_equal(np.round(x.real) + 1j * np.round(x.imag), x0), shift
        assert np.abs(x - x0).max() < 1e-3 * np.abs(x0).max()


def test_dwt3_axes_match_oracle_and_shift_hook(llr):
    """Analysis along x, y, z chained equals the oracle's single-level 3-D db2
    analysis; a shift s reads the volume at i + s (the same as analysing the
    volume rolled by -s), each axis modulo its own length."""
    from llr_hip import cudwt3 as W

    vol = (20, 13, 11)
    x = crandn(int(np.prod(vol)))
    boxes = [list(vol)]
    for a in range(3):
        b = list(boxes[-1])
        b[a] = 2 * W.sub_band_size(vol[a], 4)
        boxes.append(b)
    cur = x
    for a in range(3):
        out = np.zeros(int(np.prod(boxes[a + 1])), np.complex64)
        W.cu_dwt3_axis(cur, out, boxes[a], boxes[a + 1], a, W.DB2_LO_D, W.DB2_HI_D)
        cur = out
    ref, cdims = O.dwt3(x.astype(np.complex128), vol)
    assert list(cdims) == boxes[3]
    assert rel(cur, ref) < 2e-6
    s = (7, -4, 12)
    for a in range(3):
        dec = list(vol)
        dec[a] = boxes[a + 1][a]
        got = np.zeros(int(np.prod(dec)), np.complex64)
        W.cu_dwt3_axis(x, got, vol, dec, a, W.DB2_LO_D, W.DB2_HI_D, s)
        want = np.zeros_like(got)
        W.cu_dwt3_axis(roll3(x, vol, [-v for v in s]), want, vol, dec, a, W.DB2_LO_D, W.DB2_HI_D)
        assert np.array_equal(got, want)


def test_dwt3_axis_other_filters(llr):
    """Any even filter length: Haar and db3 (6 taps) analysis + synthesis
    reconstruct (orthogonal filters, symmetric extension)."""
    from llr_hip import cudwt3 as W

    vol = (24, 10, 6)
    x = crandn(int(np.prod(vol)))
    h = np.array([1.0, 1.0]) / np.sqrt(2)
    db3 = np.array([0.3326705529509569, 0.8068915093133388, 0.4598775021193313, -0.1350110200102546,
                    -0.0854412738822415, 0.0352262918821007])
    for rec_lo in (h, db3):
        F = rec_lo.size
        lo_d = rec_lo[::-1]
        hi_r = np.array([(-1) ** k * lo_d[k] for k in range(F)])
        hi_d = hi_r[::-1]
        dec = [2 * W.sub_band_size(vol[0], F), vol[1], vol[2]]
        r = np.zeros(int(np.prod(dec)), np.complex64)
        W.cu_dwt3_axis(x, r, vol, dec, 0, lo_d, hi_d)
        y = np.zeros_like(x)
        W.cu_idwt3_axis(y, r, vol, dec, 0, rec_lo, hi_r)
        assert rel(y, x) < 3e-6, F


def test_cycle_spinning_wavelet(llr):
    from llr_hip import recon

    vol, nv, alpha = (32, 24, 20), 2, 0.3
    N = int(np.prod(vol))
    x = crandn(nv * N)
    for s in ((0, 0, 0), (5, -7, 13)):
        g = x.copy()
        recon.wavelet_denoise_shifted(g, nv, vol, alpha, s)
        ref = np.zeros(nv * N, np.complex128)
        for v in range(nv):
            seg = roll3(x[v * N:(v + 1) * N].astype(np.complex128), vol, [-c for c in s])
            O.wavelet_denoise(seg, vol, alpha)
            ref[v * N:(v + 1) * N] = roll3(seg, vol, s)
        e = rel(g, ref)
        print(f"[cycle spinning shift {s}] rel vs oracle {e:.2e}", flush=True)
        assert e < 3e-6
    g0 = x.copy()
    recon.wavelet_denoise_shifted(g0, nv, vol, alpha, (0, 0, 0))
    g1 = x.copy()
    recon.wavelet_denoise(g1, nv, vol, alpha)
    assert np.array_equal(g0, g1)


def test_centered_fft_entry_points(llr):
    from llr_hip import signal_model as sm

    vol, nv = (32, 24, 16), 5
    N = int(np.prod(vol))
    x = crandn(nv * N)
    f = x.copy()
    sm.fft3c(f, vol, nv)
    ref = x.astype(np.complex128)
    z = np.zeros((nv, 3))
    O.signal_model(ref, vol, nv, z, z, 0)  # FP64, exact ramps: the zero-correction forward model
    assert rel(f, ref) < 2e-6
    # closed form for even sizes: prod (-1)^(n/2) fftshift(fftn(ifftshift(x)))
    v = x.reshape(nv, vol[2], vol[1], vol[0]).astype(np.complex128)
    cf = np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(v, axes=(1, 2, 3)), axes=(1, 2, 3), norm="ortho"),
                         axes=(1, 2, 3)) * np.prod([(-1) ** (n // 2) for n in vol])
    assert rel(f, cf.ravel()) < 2e-6
    fc = x.copy()
    sm.fft3c_chunked(fc, vol, nv, 2)
    assert np.array_equal(fc, f)
    b = f.copy()
    sm.ifft3c_chunked(b, vol, nv, 3)
    assert rel(b, x) < 2e-6
    b2 = f.copy()
    sm.ifft3c(b2, vol, nv)
    assert np.array_equal(b, b2)
    one = x[:N].copy()
    sm.cu_fft3(one, vol, sm.ModelDirection.Forward)
    vv = x[:N].reshape(vol[2], vol[1], vol[0]).astype(np.complex128)
    assert rel(one, np.fft.fftn(vv).ravel()) < 2e-6
    sm.cu_fft3(one, vol, sm.ModelDirection.Inverse)
    assert rel(one, x[:N]) < 2e-6


def test_recon_with_image_corrections(llr):
    """lin_img_corr (signal_model.rs:34-79) through the resident reconstruction,
    against the CPU FP32 reference path with the same corrections."""
    from llr_hip import recon

    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=3))
    cfg = ds["cfg"]
    ic = rng.uniform(-1.5, 1.5, (ds["nq"], 3)).astype(np.float32)
    for hybrid in (True, False):
        r = recon.Recon(ds["vol"], ds["nq"], ds["shifts"], blk=cfg["blk"], lambda_factor=cfg["lambda_factor"],
                        wavelet_alpha=cfg["wavelet_alpha"], subspace=ds["subspace"], img_corr=ic, hybrid=hybrid)
        r.set_measurements(ds["meas"])
        h = r.run()
        img = r.image()
        r.close()
        x32, h32, _, _ = O.recon(ds["meas"], ds["vol"], ds["nq"], cfg["blk"], ds["shifts"], llr_mode=1,
                                 lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"],
                                 subspace=ds["subspace"], img_corr=ic)
        x0 = O.recon(ds["meas"], ds["vol"], ds["nq"], cfg["blk"], ds["shifts"], llr_mode=1,
                     lambda_factor=cfg["lambda_factor"], wavelet_alpha=cfg["wavelet_alpha"],
                     subspace=ds["subspace"])[0]
        e = H.nrmse(img, x32)
        print(f"[recon with image corrections, hybrid {hybrid}] NRMSE vs cpu32 {e:.3e} (vs zero corrections "
              f"{H.nrmse(img, x0):.3e})", flush=True)
        assert e <= 1e-4 and H.nrmse(img, x0) > 1e-2
        assert np.allclose(h, h32, rtol=1e-3)
