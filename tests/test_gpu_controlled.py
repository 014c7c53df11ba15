"""Controlled evaluation (nfgs_amd.controlled.evaluate) on the GPU.

The reference's own evaluation test (tests/test_controlled_evaluation.py:35-50:
the anisotropic polynomial bundle with the example configuration, held here
as tests/golden/evaluation_example.json) and every metric against the same
evaluation assembled from the oracle (FD2 curl/divergence, td_metrics, the
outward Gaussian field, outward_errors).  The device cci_errors
(nfgs_outward_errors) is also checked against the oracle's outward_errors
on random planes, windows and cutoffs, and on a zero target.
"""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import nfgs_oracle as O
from tests.test_controlled_contract import _bundle

pytestmark = pytest.mark.gpu

CONFIG = Path(__file__).resolve().parent / "golden" / "evaluation_example.json"


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _noisy_boundary(root):
    from nfgs_amd.controlled import load
    obs = load(root / "boundary.npz", True)
    obs["Bz"] = obs["Bz"] + 0.1 * np.outer(np.sin(obs["y"]), np.cos(obs["x"]))
    np.savez(root / "boundary.npz", **obs)


def test_identity_and_actual_observation_target(tmp_path):
    _announce("controlled evaluation identity / noisy boundary")
    from nfgs_amd.controlled import evaluate
    _bundle(tmp_path)
    args = (tmp_path / "field.npz", tmp_path / "field.npz", tmp_path / "boundary.npz", CONFIG)
    result = evaluate(*args)
    m = result["metrics"]
    assert m["C_vec"] == pytest.approx(1.0, abs=1e-12)
    assert m["Eprime_n"] == pytest.approx(1.0, abs=1e-12)
    assert m["E_rel_Bgt"] == pytest.approx(0.0, abs=1e-12)
    assert result["current_operator"] == "common_fd2_float64_full_volume_then_crop"
    assert len(result["source_sha256"]) == 4
    _noisy_boundary(tmp_path)
    noisy = evaluate(*args)["metrics"]
    assert noisy["E_rel_Bgt_L2"] > 0
    assert noisy["clean_reference_CCI"]["E_rel_Bgt_L2"] == pytest.approx(0.0, abs=1e-12)


def _oracle_metrics(ref, field, obs, cfg):
    cci_errors = O.outward_errors
    box = tuple(slice(*s) for s in cfg["scoring_slices_zyx"])
    h = tuple(float(np.diff(ref[k])[0]) for k in ("x", "y", "z"))
    j = O.fd2_curl(field, h)
    div = O.fd2_divergence(field, h)
    m = O.td_metrics(ref, field, j, div, box, 3.0 / sum(1.0 / v for v in h))

    def split(f0):
        o = O.OracleNFGS(len(ref["x"]), len(ref["y"]), 8, *h, dtype=np.float64,
                         pad_calculation_domain=False, source_zero_mode_policy="project",
                         d0_zero_mode_policy="project")
        o.set_bottom_vector_boundary(f0["Bx"], f0["By"], f0["Bz"])
        return o.gaussian_above_bottom_field()

    cand = split({k: field[k][0] for k in ("Bx", "By", "Bz")})
    cuts = cfg["CCI_cutoff_fractions"]
    m.update(cci_errors(split(obs), cand, box, cuts))
    m["clean_reference_CCI"] = cci_errors(split({k: ref[k][0] for k in ("Bx", "By", "Bz")}),
                                          cand, box, cuts)
    _, sy, sx = box
    m["bottom_Bz_RMSE_to_input"] = float(np.sqrt(np.mean((field["Bz"][0, sy, sx] -
                                                          obs["Bz"][sy, sx]) ** 2)))
    jobs = O.fd2_curl({k: np.repeat(obs[k][None], 3, axis=0) for k in ("Bx", "By", "Bz")}, h)
    m["bottom_Jz_RMSE_to_input_fd2"] = float(np.sqrt(np.mean((j["Jz"][0, sy, sx] -
                                                              jobs["Jz"][0][sy, sx]) ** 2)))
    return m


def _close(got, want, path=""):
    if isinstance(want, dict):
        assert set(got) == set(want), path
        for k in want:
            _close(got[k], want[k], f"{path}.{k}")
    elif want is None or isinstance(want, (int, np.integer)):
        assert got == want, path
    else:
        assert got == pytest.approx(want, rel=1e-10, abs=1e-12), path


def test_every_metric_matches_the_oracle_evaluation(tmp_path):
    _announce("controlled evaluation vs oracle")
    import json
    from nfgs_amd.controlled import evaluate, load
    _bundle(tmp_path)
    rng = np.random.default_rng(4)
    cand = load(tmp_path / "field.npz")
    for k in ("Bx", "By", "Bz"):
        cand[k] = cand[k] + 0.05 * rng.standard_normal(cand[k].shape)
    np.savez(tmp_path / "cand.npz", **cand)
    _noisy_boundary(tmp_path)
    cfg = json.loads(CONFIG.read_text())
    cfg["scoring_slices_zyx"] = [[1, 4, 1], [1, 6, 1], [2, 8, 1]]
    (tmp_path / "cfg.json").write_text(json.dumps(cfg))
    got = evaluate(tmp_path / "field.npz", tmp_path / "cand.npz", tmp_path / "boundary.npz",
                   tmp_path / "cfg.json")["metrics"]
    want = _oracle_metrics(load(tmp_path / "field.npz"), cand,
                           load(tmp_path / "boundary.npz", True), cfg)
    _close(got, want)


def test_bad_configs_are_rejected(tmp_path):
    _announce("controlled evaluation errors")
    import json
    from nfgs_amd.controlled import evaluate
    _bundle(tmp_path)
    cfg = json.loads(CONFIG.read_text())
    bad = dict(cfg, scoring_slices_zyx=[[0, 6, 1], [0, 7, 1], [0, 9, 1]])
    (tmp_path / "bad.json").write_text(json.dumps(bad))
    f = tmp_path / "field.npz"
    with pytest.raises(ValueError, match="outside the field"):
        evaluate(f, f, tmp_path / "boundary.npz", tmp_path / "bad.json")
    (tmp_path / "v2.json").write_text(json.dumps(dict(cfg, schema_version=2)))
    with pytest.raises(ValueError, match="Unsupported evaluation config"):
        evaluate(f, f, tmp_path / "boundary.npz", tmp_path / "v2.json")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_device_cci_errors_match_the_oracle(seed):
    _announce(f"device cci_errors seed {seed}")
    from nfgs_amd.controlled import cci_errors
    rng = np.random.default_rng(seed)
    ny, nx = 37 + 5 * seed, 53 - 3 * seed
    b = rng.standard_normal((3, ny, nx)) * np.exp(rng.standard_normal((1, ny, nx)))
    c = b + 0.1 * rng.standard_normal(b.shape)
    cuts = [1e-8, 0.01, 0.05, 0.3, 0.9]
    for box in ((slice(None),) * 3, (slice(None), slice(3, ny - 4), slice(1, nx - 7))):
        got, want = cci_errors(b, c, box, cuts), O.outward_errors(b, c, box, cuts)
        _close(got, want)
    zero = cci_errors(np.zeros((3, 7, 9)), np.ones((3, 7, 9)), (slice(None),) * 3, [1e-8, .5])
    assert zero == O.outward_errors(np.zeros((3, 7, 9)), np.ones((3, 7, 9)), (slice(None),) * 3,
                                    [1e-8, .5])
    assert zero["E_rel_Bgt_L2"] is None and zero["E_rel_Bgt"] is None
