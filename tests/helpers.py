"""Shared test helpers: data sets built with the oracle's FP64 forward model."""
import numpy as np

import oracle as O
from llr_hip import phantom


def oracle_forward(x, vol, nq):
    """Forward signal model (zero corrections) via the FP64 oracle, in place on complex64."""
    y = x.astype(np.complex128)
    z = np.zeros((nq, 3))
    O.signal_model(y, vol, nq, z, z, 0)
    x[:] = y.astype(np.complex64)


def nrmse(a, b):
    return float(np.linalg.norm(np.asarray(a, np.complex128) - np.asarray(b, np.complex128)) /
                 np.linalg.norm(np.asarray(b, np.complex128)))


SMALL_SUBSPACE = dict(vol=(32, 32, 16), n_b0=1, dirs="repulsion", n_dirs=12, bval=1000.0, R=4.0, n_iter=6,
                      blk=(8, 8, 8), lambda_factor=0.02, wavelet_alpha=0.005, srank=7, n_ell=8, seeds=(21, 22, 23, 24))

SMALL_PLAIN = dict(vol=(32, 32, 32), n_b0=1, dirs="six", n_dirs=6, bval=1000.0, R=4.0, n_iter=8, blk=(8, 8, 8),
                   lambda_factor=0.02, wavelet_alpha=0.0, srank=0, n_ell=8, seeds=(31, 32, 33, 34))


def make(preset):
    return phantom.dataset(preset, oracle_forward)


def nrmse_big(a, b, chunk=1 << 24):
    """nrmse for arrays of hundreds of millions of samples, in FP64 chunks."""
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    num = den = 0.0
    for i in range(0, a.size, chunk):
        bb = b[i:i + chunk].astype(np.complex128)
        d = a[i:i + chunk].astype(np.complex128) - bb
        num += float(np.vdot(d, d).real)
        den += float(np.vdot(bb, bb).real)
    return float(np.sqrt(num / den))
