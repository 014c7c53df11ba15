"""The NPZ reconstruction entry (nfgs_amd.reconstruct_npz) on the GPU.

The formats of the reference's examples/reconstruct_npz.py:19-67 (input
Bx/By/Bz, dx/dy/dz, Nz; output six fields plus a JSON summary) and the
example's solve settings.  The golden fp32_rho05 case was written by the
reference with exactly those settings at max_iter = 30, so the command line
run on its boundary reproduces the reference's output fields.
"""
import json
import os
import sys

import numpy as np
import pytest
import torch

from tests.golden_cases import case_files, load, rel_l2

pytestmark = pytest.mark.gpu

CASE = [p for p in case_files() if os.path.basename(p) == "fp32_rho05.npz"][0]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_command_line_reproduces_the_reference_example(tmp_path):
    _announce("reconstruct_npz fp32_rho05")
    from nfgs_amd.reconstruct_npz import main
    g, meta = load(CASE)
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    np.savez(tmp_path / "boundary.npz", Bx=g["in_bx"], By=g["in_by"], Bz=g["in_bz"], dx=1.0,
             dy=1.0, dz=1.0, Nz=nz)
    out = tmp_path / "run" / "rec.npz"
    assert main([str(tmp_path / "boundary.npz"), "--output", str(out), "--rho", str(c["rho"]),
                 "--max-iter", str(c["iterate"]["max_iter"])]) == 0
    with np.load(out) as d:
        fields = {k: d[k] for k in d.files}
    assert sorted(fields) == ["Bx", "By", "Bz", "Jx", "Jy", "Jz"]
    for k, v in fields.items():
        assert v.shape == (nz, nyo, nxo) == g["out_" + k].shape, k
    summary = json.loads(out.with_suffix(".json").read_text())
    assert set(summary) == {"rho", "best_iter", "best_sigma", "stop_reason"}
    assert summary["rho"] == c["rho"]
    assert summary["stop_reason"] == meta["stop_reason"]
    assert summary["best_sigma"] == pytest.approx(meta["best_sigma"], rel=2e-4)
    # the golden FP32 run is fixed-length (sigma falls at all 30 steps)
    assert summary["best_iter"] == meta["best_iter"] == c["iterate"]["max_iter"]
    for k in ("Bx", "By", "Bz"):
        assert rel_l2(fields[k], g["out_" + k]) < 1e-4, k


def test_outputs_equal_a_direct_solve(tmp_path):
    _announce("reconstruct_npz vs direct solve")
    from nfgs_amd.reconstruct_npz import main, read_boundary, reconstruct
    g, meta = load(CASE)
    nz = meta["case"]["observed"][2]
    np.savez(tmp_path / "b.npz", Bx=g["in_bx"], By=g["in_by"], Bz=g["in_bz"], dx=1.0, dy=1.0,
             dz=1.0, Nz=nz)
    main([str(tmp_path / "b.npz"), "--output", str(tmp_path / "o.npz"), "--max-iter", "5"])
    bx, by, bz, nz2, h = read_boundary(tmp_path / "b.npz")
    assert nz2 == nz and h == (1.0, 1.0, 1.0)
    s, hist = reconstruct(bx, by, bz, nz2, h, rho=0.5, max_iter=5, verbose=False)
    direct = s.fields_to_numpy(crop_to_observed_fov=True)
    with np.load(tmp_path / "o.npz") as d:
        for k in direct:
            np.testing.assert_array_equal(d[k], direct[k])


def test_rejects_bad_arguments(tmp_path):
    _announce("reconstruct_npz errors")
    from nfgs_amd.reconstruct_npz import main
    np.savez(tmp_path / "b.npz", Bx=np.zeros((4, 5)), By=np.zeros((4, 5)), Bz=np.zeros((5, 4)),
             dx=1.0, dy=1.0, dz=1.0, Nz=4)
    with pytest.raises(SystemExit):
        main([str(tmp_path / "b.npz"), "--rho", "1.5"])
    with pytest.raises(SystemExit):
        main([str(tmp_path / "b.npz"), "--device", "cpu"])
    with pytest.raises(ValueError, match="one shape"):
        main([str(tmp_path / "b.npz")])
