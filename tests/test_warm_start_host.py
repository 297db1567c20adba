"""Host-side logic of the learned warm start (no GPU): the coefficient <->
network-channel maps. The network must see the same physical quantity in each
channel whatever the b-table, or a network trained on one set of directions
would meet rotated channels on another."""
import numpy as np

from llr_hip import phantom
from llr_hip import warm_start as W


def table(n_b0, n_dirs, seed, bval=3000.0):
    dirs = phantom.repulsion_directions(n_dirs, seed, iters=100)
    return phantom.b_table(n_b0, dirs, bval)


def test_maps_are_inverse():
    bv, bq = table(1, 30, 7)
    S = phantom.loglinear_subspace(bv, bq, 7)
    to, fr = W.tensor_mixing(bv, bq, S)
    assert to.shape == (7, 7) and fr.shape == (7, 7)
    assert np.allclose(to.astype(np.float64) @ fr, np.eye(7), atol=1e-5)
    assert np.allclose(fr.astype(np.float64) @ to, np.eye(7), atol=1e-5)


def test_channels_do_not_depend_on_the_direction_set():
    # a voxel whose signal lies in the design space: x = A beta (b in units of 1000)
    rng = np.random.default_rng(3)
    beta = rng.standard_normal(7) + 1j * rng.standard_normal(7)
    got = []
    for n_b0, n_dirs, seed in [(1, 30, 11), (7, 60, 12), (2, 12, 13)]:
        bv, bq = table(n_b0, n_dirs, seed)
        A = phantom.loglinear_design(bv / 1000.0, bq)
        x = A @ beta
        S = phantom.loglinear_subspace(bv, bq, 7).astype(np.complex128)
        c = S.T @ x  # the subspace coefficients the reconstruction works on (S is real)
        to, _ = W.tensor_mixing(bv, bq, S)
        got.append(to.astype(np.float64) @ c)
    for g in got:
        assert np.allclose(g, beta, rtol=1e-5, atol=1e-5)
    # while the raw subspace coefficients of the same voxel differ between tables
    bv1, bq1 = table(1, 30, 11)
    bv2, bq2 = table(7, 60, 12)
    c1 = phantom.loglinear_subspace(bv1, bq1, 7).T @ (phantom.loglinear_design(bv1 / 1000.0, bq1) @ beta)
    c2 = phantom.loglinear_subspace(bv2, bq2, 7).T @ (phantom.loglinear_design(bv2 / 1000.0, bq2) @ beta)
    assert not np.allclose(c1, c2, rtol=1e-2)


def test_rank_deficient_subspace_uses_identity_maps():
    bv, bq = table(1, 30, 7)
    S = phantom.loglinear_subspace(bv, bq, 5)
    to, fr = W.tensor_mixing(bv, bq, S)
    assert np.array_equal(to, np.eye(5, dtype=np.float32)) and np.array_equal(fr, np.eye(5, dtype=np.float32))


def test_member_config_seeds():
    cfg = W.member_config(1234)
    assert cfg["seeds"] == (1234, 2234, 3234, 4234)
    assert cfg["vol"] == (128, 128, 96) and cfg["n_b0"] + cfg["n_dirs"] == 31
