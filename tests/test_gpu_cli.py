"""llr-recon end to end (bin/llr-recon.rs:6-25, lib.rs:113-363): a data set of
per-volume k-space .cfl files, a b-table with a b0 volume, a (ny, nz, nq)
phase-encode mask and a subspace matrix for the diffusion-weighted volumes,
reconstructed by the executable in the reference's rank mode. Its output
volumes must equal, bit for bit, the library's reconstruction of the same
measurements with the block shifts it wrote (shifts.txt), its
error_history.json must hold that run's residuals, and the image must match
the CPU FP32 reference path within 1e-4 relative NRMSE."""
import json
import os
import subprocess

import numpy as np
import pytest

import helpers as H
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "llr-reco_amd", "bin", "llr-recon")


def test_llr_recon_cli_end_to_end(llr, tmp_path):
    from llr_hip import data_import as di
    from llr_hip import phantom, recon

    ds = H.make(dict(H.SMALL_SUBSPACE, n_iter=3))
    vol, nq = ds["vol"], ds["nq"]
    N = int(np.prod(vol))
    kdir = tmp_path / "ksp"
    kdir.mkdir()
    for q in range(nq):
        di.write_cfl(str(kdir / f"ksp_{q:02d}"), ds["meas"][q * N:(q + 1) * N], list(vol))
    rows = np.concatenate([ds["bvals"][:, None], ds["bvecs"]], axis=1)
    (tmp_path / "btab.txt").write_text("\n".join(" ".join(f"{v:.9g}" for v in r) for r in rows) + "\n")
    di.write_cfl(str(tmp_path / "mask"), ds["mask"].ravel(order="F"), [vol[1], vol[2], nq])
    dti = np.flatnonzero(ds["bvals"] > 0)
    sub = phantom.loglinear_subspace(ds["bvals"][dti], ds["bvecs"][dti], 7)
    di.write_cfl(str(tmp_path / "subspace"), sub.ravel(order="F"), list(sub.shape))
    (tmp_path / "params.toml").write_text(
        "llr_block_size = [8, 8, 8]\nfft_max_batch_size = 200\ncas_mat_max_batch_size = 200\nllr_block_rank = 4\n"
        f"wavelet_shrink_factor = 0.005\nn_iter = 3\nsubspace_projection_mat = \"{tmp_path / 'subspace.cfl'}\"\n"
        "shift_seed = 5\n")
    out = tmp_path / "out"
    out.mkdir()
    cmd = [EXE, str(tmp_path / "params"), str(kdir), "ksp_*", str(out), "img", str(tmp_path / "btab"), "0",
           str(tmp_path / "mask"), str(nq)]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    shifts = np.loadtxt(out / "shifts.txt", dtype=np.uint64).reshape(-1, 3)
    hist_cli = json.loads((out / "error_history.json").read_text())
    width = len(str(nq))
    img_cli = np.concatenate([di.read_cfl(str(out / f"img{q:0{width}d}"))[0] for q in dti])
    # the library on the same measurements (the DTI volumes, masked as the CLI does)
    meas = np.concatenate([ds["meas"][q * N:(q + 1) * N] for q in dti])
    lib = recon.Recon(vol, len(dti), shifts, blk=(8, 8, 8), llr_mode="rank", rank=4, wavelet_alpha=0.005,
                      subspace=sub)
    lib.set_measurements(meas)
    h = lib.run()
    img = lib.image()
    lib.close()
    x32, h32, _, _ = O.recon(meas, vol, len(dti), (8, 8, 8), shifts, llr_mode=0, rank=4, wavelet_alpha=0.005,
                             subspace=sub)
    e = H.nrmse(img_cli, x32)
    print(f"[llr-recon] {len(dti)} DTI volumes, shifts {shifts.tolist()}; NRMSE vs cpu32 {e:.3e}; residuals "
          f"{hist_cli}", flush=True)
    assert np.array_equal(img_cli, img)
    assert hist_cli == [float(v) for v in h]
    assert e <= 1e-4
    assert np.allclose(h, h32, rtol=1e-3)
