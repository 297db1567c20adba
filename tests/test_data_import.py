"""Data ingest and egress (host-only C ABI calls, no GPU): BART .cfl/.hdr,
the DTI data-set builder (data_import.rs:13-273), b-table and b0 mask,
LlrReconParams TOML (lib.rs:47-80, bin/llr-reco-params.rs:14-23),
error_history.json (lib.rs:345-351), and the llr-recon / llr-reco-params
executables up to the point where they would use the GPU.

The .cfl format is pinned by an independent numpy reader/writer of BART's
public layout (text header "# Dimensions" + sizes, raw column-major complex64
little endian)."""
import json
import os
import subprocess

import numpy as np
import pytest

from llr_hip import _lib
from llr_hip import data_import as di

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "llr-reco_amd", "bin")
rng = np.random.default_rng(99)


def crandn(n):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)


def np_write_cfl(base, data, dims):
    with open(base + ".hdr", "w") as f:
        f.write("# Dimensions\n" + " ".join(str(d) for d in dims) + "\n")
    np.asarray(data, np.complex64).tofile(base + ".cfl")


def np_read_cfl(base):
    lines = [ln for ln in open(base + ".hdr").read().splitlines() if ln and not ln.startswith("#")]
    dims = [int(x) for x in lines[0].split()]
    return np.fromfile(base + ".cfl", np.complex64), dims


# ------------------------------------------------------------------ cfl
def test_cfl_round_trip_and_format(tmp_path):
    base = str(tmp_path / "vol")
    x = crandn(6 * 5 * 4)
    di.write_cfl(base, x, [6, 5, 4])
    hdr = open(base + ".hdr").read()
    assert hdr == "# Dimensions\n6 5 4 " + "1 " * 13 + "\n"
    y, dims = np_read_cfl(base)
    assert dims[:3] == [6, 5, 4] and all(d == 1 for d in dims[3:]) and len(dims) == 16
    assert np.array_equal(y, x)
    # and a file written by the independent writer reads back through the library
    np_write_cfl(base + "2", x, [6, 1, 5, 4])
    z, d2 = di.read_cfl(base + "2.cfl")
    assert d2 == [6, 1, 5, 4] and np.array_equal(z, x)
    with pytest.raises(_lib.LlrError):
        di.write_cfl(base, x, [6, 5, 5])  # length mismatch


def test_cfl_missing_and_short_files(tmp_path):
    base = str(tmp_path / "v")
    with pytest.raises(_lib.LlrError, match="failed to open header"):
        di.read_cfl(base)
    np_write_cfl(base, crandn(10), [4, 3])  # 12 samples promised, 10 stored
    with pytest.raises(_lib.LlrError, match="short read"):
        di.read_cfl(base)


# ------------------------------------------------------------------ b-table
def test_b_table_and_b0_mask(tmp_path):
    p = tmp_path / "btab.txt"
    rows = [[0, 0, 0, 0], [1000, 1, 0, 0], [1000, 0, 0.6, 0.8], [5, 1e-3, 0, 0]]
    p.write_text("\n".join(" ".join(str(v) for v in r) for r in rows) + "\n")
    bv, g = di.read_b_table(str(tmp_path / "btab"))  # extension added as with_extension("txt")
    assert np.allclose(bv, [0, 1000, 1000, 5]) and np.allclose(g, np.array(rows)[:, 1:])
    assert list(di.b0_mask(g, 0.0)) == [True, False, False, False]
    assert list(di.b0_mask(g, 1e-2)) == [True, False, False, True]
    with pytest.raises(_lib.LlrError, match="non-negative"):
        di.b0_mask(g, -1.0)
    p.write_text("0 0 0 0 1000 1 0\n")
    with pytest.raises(_lib.LlrError, match="divisible by 4"):
        di.read_b_table(str(p))
    p.write_text("0 0 0 zero\n")
    with pytest.raises(_lib.LlrError, match="failed to parse b-table at character: zero"):
        di.read_b_table(str(p))


# ------------------------------------------------------------------ data set
def make_dir(tmp_path, n=5, vol=(8, 6, 4), b0=(0, 3)):
    d = tmp_path / "cfl"
    d.mkdir()
    vols = []
    for i in range(n):
        x = crandn(int(np.prod(vol)))
        dims = list(vol) if i != 2 else [vol[0], 1, vol[1], vol[2]]  # singleton dims are dropped
        np_write_cfl(str(d / f"k{i:02d}"), x, dims)
        vols.append(x)
    np_write_cfl(str(d / "other"), crandn(4), [4])  # does not match the pattern
    rows = [[0, 0, 0, 0] if i in b0 else [1000, *(np.eye(3)[i % 3])] for i in range(n)]
    (tmp_path / "bt.txt").write_text("\n".join(" ".join(str(v) for v in r) for r in rows))
    return d, vols


def test_dataset_builder(tmp_path):
    d, vols = make_dir(tmp_path)
    ds = di.DtiDataSetBuilder.from_cfl_volume_dir(str(d), "k*", 5).with_b_table(str(tmp_path / "bt"), 0.0)
    info = ds.info()
    assert info == dict(vol_dims=(8, 6, 4), total_volumes=5, n_dti=3, n_b0=2)
    assert [os.path.basename(ds.file(i)) for i in range(5)] == [f"k{i:02d}.cfl" for i in range(5)]
    assert ds.dti_indices() == [1, 2, 4] and ds.b0_indices() == [0, 3]
    assert np.array_equal(ds.load_dti_set(), np.concatenate([vols[i] for i in (1, 2, 4)]))
    assert np.array_equal(ds.load_b0_set(), np.concatenate([vols[i] for i in (0, 3)]))
    assert np.array_equal(ds.load_all(), np.concatenate(vols))
    bv, g = ds.b_table()
    assert list(bv) == [0, 1000, 1000, 0, 1000]
    mask = (rng.random((6, 4, 5)) < 0.5).astype(np.complex64)
    np_write_cfl(str(tmp_path / "mask"), mask.ravel(order="F"), [6, 4, 5])
    ds.with_sample_mask(str(tmp_path / "mask"))
    flat = mask.ravel(order="F").reshape(5, 24)
    assert np.array_equal(ds.dti_sample_mask(), flat[[1, 2, 4]].ravel())
    assert np.array_equal(ds.b0_sample_mask(), flat[[0, 3]].ravel())
    ds.close()


def test_dataset_builder_errors(tmp_path):
    d, _ = make_dir(tmp_path)
    with pytest.raises(_lib.LlrError, match="unexpected number of files"):
        di.DtiDataSetBuilder.from_cfl_volume_dir(str(d), "k*", 4)
    with pytest.raises(_lib.LlrError, match="no files found"):
        di.DtiDataSetBuilder.from_cfl_volume_dir(str(d), "nothing*")
    np_write_cfl(str(d / "k99"), crandn(8 * 6 * 5), [8, 6, 5])
    with pytest.raises(_lib.LlrError, match="mismatch between first volume dims"):
        di.DtiDataSetBuilder.from_cfl_volume_dir(str(d), "k*")
    os.remove(str(d / "k99.cfl"))
    os.remove(str(d / "k99.hdr"))
    ds = di.DtiDataSetBuilder.from_cfl_volume_dir(str(d), "k*")
    with pytest.raises(_lib.LlrError, match="you must load a b-table first"):
        ds.load_dti_set()
    (tmp_path / "bt4.txt").write_text("0 0 0 0\n" * 4)
    with pytest.raises(_lib.LlrError, match="expected 5 bvecs, received 4"):
        ds.with_b_table(str(tmp_path / "bt4"))
    ds.with_b_table(str(tmp_path / "bt"))
    np_write_cfl(str(tmp_path / "m"), np.ones(5 * 4 * 5, np.complex64), [5, 4, 5])
    with pytest.raises(_lib.LlrError, match="phase encoding dimension mismatch"):
        ds.with_sample_mask(str(tmp_path / "m"))
    np_write_cfl(str(tmp_path / "m2"), np.ones(6 * 4 * 3, np.complex64), [6, 4, 3])
    with pytest.raises(_lib.LlrError, match="sample mask volumes"):
        ds.with_sample_mask(str(tmp_path / "m2"))


# ------------------------------------------------------------------ params
DEFAULT_TOML = """llr_block_size = [32, 32, 32]
fft_max_batch_size = 200
cas_mat_max_batch_size = 200
llr_block_rank = 6
wavelet_shrink_factor = 0.005
n_iter = 60
subspace_projection_mat = "/path/to/matrix.cfl"

"""


def test_default_params_file(tmp_path):
    """LlrReconParams::default (lib.rs:67-80) as toml::to_string writes it (fields
    in declaration order, the None max_rand_shift left out)."""
    di.write_recon_params(str(tmp_path / "p"))
    assert (tmp_path / "p.toml").read_text() == DEFAULT_TOML
    p = di.read_recon_params(str(tmp_path / "p"))
    assert p["llr_block_size"] == (32, 32, 32) and p["max_rand_shift"] is None
    assert (p["fft_max_batch_size"], p["cas_mat_max_batch_size"], p["llr_block_rank"], p["n_iter"]) == (200, 200, 6, 60)
    assert abs(p["wavelet_shrink_factor"] - 0.005) < 1e-9 and p["subspace_projection_mat"] == "/path/to/matrix.cfl"
    assert p["llr_mode"] == "rank"


def test_params_optional_extension_and_errors(tmp_path):
    f = tmp_path / "q.toml"
    f.write_text("# comment\nllr_block_size = [8, 8, 8]\nmax_rand_shift = [4, 4, 2]\nfft_max_batch_size = 10\n"
                 "cas_mat_max_batch_size = 20\nllr_block_rank = 3 # trailing\nwavelet_shrink_factor = 1e-3\n"
                 "n_iter = 5\nllr_mode = \"svt\"\nsvt_lambda_factor = 0.05\nstorage = \"fp16\"\n"
                 "precision = \"mixed\"\nshift_seed = 7\nsomething_else = \"ignored\"\n")
    p = di.read_recon_params(str(f))
    assert p["max_rand_shift"] == (4, 4, 2) and p["subspace_projection_mat"] is None
    assert (p["llr_mode"], p["storage"], p["precision"], p["shift_seed"]) == ("svt", "fp16", "mixed", 7)
    assert abs(p["svt_lambda_factor"] - 0.05) < 1e-12 and abs(p["wavelet_shrink_factor"] - 1e-3) < 1e-9
    f.write_text("llr_block_size = [8, 8, 8]\nfft_max_batch_size = 10\ncas_mat_max_batch_size = 20\n"
                 "wavelet_shrink_factor = 0.1\nn_iter = 5\n")
    with pytest.raises(_lib.LlrError, match="missing field `llr_block_rank`"):
        di.read_recon_params(str(f))
    f.write_text("llr_block_size = [8, 8]\nfft_max_batch_size = 10\ncas_mat_max_batch_size = 20\n"
                 "llr_block_rank = 3\nwavelet_shrink_factor = 0.1\nn_iter = 5\n")
    with pytest.raises(_lib.LlrError, match="llr_block_size: expected an array of 3 integers"):
        di.read_recon_params(str(f))
    with pytest.raises(_lib.LlrError, match="failed to open config file"):
        di.read_recon_params(str(tmp_path / "absent"))


def test_error_history_json(tmp_path):
    h = [0.1, 1.0, 1e-5, 0.0070996, 2.5e-12, 123456.75]
    p = str(tmp_path / "error_history.json")
    di.write_error_history(p, h)
    text = open(p).read()
    assert text == "[\n  0.1,\n  1.0,\n  1e-5,\n  0.0070996,\n  2.5e-12,\n  123456.75\n]\n"
    assert json.loads(text) == h
    di.write_error_history(p, np.array([1 / 3, 2 / 3]))
    assert json.loads(open(p).read()) == [1 / 3, 2 / 3]
    di.write_error_history(p, [])
    assert open(p).read() == "[]\n"


# ------------------------------------------------------------------ executables
def run(args, **kw):
    return subprocess.run(args, capture_output=True, text=True, timeout=60, **kw)


def test_llr_reco_params_executable(tmp_path):
    r = run([os.path.join(BIN, "llr-reco-params"), str(tmp_path / "params")])
    assert r.returncode == 0, r.stderr
    assert "wrote default parameter file to" in r.stdout
    assert (tmp_path / "params.toml").read_text() == DEFAULT_TOML


def test_llr_recon_arguments_and_input_errors(tmp_path):
    exe = os.path.join(BIN, "llr-recon")
    r = run([exe, "--help"])
    assert r.returncode == 0 and "usage: llr-recon" in r.stdout
    r = run([exe, "a", "b"])
    assert r.returncode == 2 and "usage" in r.stderr
    d, _ = make_dir(tmp_path)
    out = tmp_path / "out"
    out.mkdir()
    args = [exe, str(tmp_path / "params"), str(d), "k*", str(out), "img", str(tmp_path / "bt")]
    r = run(args)  # no params file yet
    assert r.returncode == 1 and "failed to open config file" in r.stderr
    di.write_recon_params(str(tmp_path / "params"), dict(llr_block_size=(4, 4, 4), llr_block_rank=2, n_iter=2,
                                                        subspace_projection_mat=None))
    r = run(args + ["0", str(tmp_path / "nomask"), "4"])
    assert r.returncode == 1 and "unexpected number of files" in r.stderr
    r = run(args + ["0", str(tmp_path / "nomask")])
    assert r.returncode == 1 and "failed to open header" in r.stderr
    r = run(args + ["x"])
    assert r.returncode == 1 and "not a number" in r.stderr


def test_read_affine_translation(tmp_path):
    """signal_model.rs:119-176: the 'Parameters:' line of an affine text file (ITK
    style) holds a 3x3 matrix and a translation; the translation is returned
    times the multiplier."""
    from llr_hip import signal_model as sm

    files = []
    for i, t in enumerate([(1.5, -2.0, 0.25), (0.0, 3.0, -1.0)]):
        f = tmp_path / f"aff{i}.txt"
        f.write_text("#Insight Transform File V1.0\nTransform: AffineTransform_double_3_3\n"
                     f"Parameters: 1 0 0 0 1 0 0 0 1 {t[0]} {t[1]} {t[2]}\nFixedParameters: 0 0 0\n")
        files.append(str(f))
    out = sm.read_affine_translation(files, multiplier=2.0)
    assert np.allclose(out, [[3.0, -4.0, 0.5], [0.0, 6.0, -2.0]])
    out = sm.read_affine_translation_range(range(2), lambda i: str(tmp_path / f"aff{i}.txt"))
    assert np.allclose(out[1], [0.0, 3.0, -1.0])
    (tmp_path / "bad.txt").write_text("Parameters: 1 2 3\n")
    with pytest.raises(_lib.LlrError, match="expected 12 matrix entries"):
        sm.read_affine_translation([str(tmp_path / "bad.txt")])
    (tmp_path / "none.txt").write_text("nothing here\n")
    with pytest.raises(_lib.LlrError, match="failed to find Parameter field"):
        sm.read_affine_translation([str(tmp_path / "none.txt")])
