"""GPU common scoring (nfgs_amd.scores) against the oracle (FP64)."""
import sys

import numpy as np
import pytest
import torch

from oracle import nfgs_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _fields(shape, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(shape) for k in ("Bx", "By", "Bz", "Jx", "Jy", "Jz")}


@pytest.mark.parametrize("shape", [(5, 7, 9), (2, 6, 3), (1, 4, 5), (9, 17, 33)])
def test_curl_divergence_match_oracle(shape):
    from nfgs_amd import scores as S
    _announce(f"fd2 curl/div {shape}")
    f = _fields(shape, sum(shape))
    sp = (0.4, 1.3, 0.7)
    j, jo = S.curl(f, sp), O.fd2_curl(f, sp)
    for k in ("Jx", "Jy", "Jz"):
        np.testing.assert_allclose(j[k], jo[k], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(S.divergence(f, sp), O.fd2_divergence(f, sp), rtol=1e-13,
                               atol=1e-13)


def test_reference_known_answer_on_the_gpu():
    from nfgs_amd import scores as S
    _announce("fd2 known answer")
    x = np.linspace(-1.0, 1.0, 9)
    y = np.linspace(-2.0, 2.0, 7)
    z = np.linspace(0.0, 3.0, 5)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    f = {"Bx": -yy, "By": xx, "Bz": 1.0 + xx * yy}
    j = S.curl(f, (0.25, 4.0 / 6.0, 0.75))
    np.testing.assert_allclose(j["Jx"], xx, atol=2e-14)
    np.testing.assert_allclose(j["Jy"], -yy, atol=2e-14)
    np.testing.assert_allclose(j["Jz"], 2.0, atol=2e-14)
    np.testing.assert_allclose(S.divergence(f, (0.25, 4.0 / 6.0, 0.75)), 0.0, atol=2e-14)


@pytest.mark.parametrize("shape,box", [((6, 8, 10), (slice(1, 5), slice(0, 8), slice(2, 9))),
                                       ((33, 40, 37), (slice(0, 24), slice(5, 35), slice(4, 33)))])
def test_metrics_match_oracle(shape, box):
    from nfgs_amd import scores as S
    _announce(f"scores {shape}")
    f = _fields(shape, 3)
    ref = {k: v + 0.1 * np.sin(v) for k, v in _fields(shape, 4).items()}
    m, mo = S.compute_reference_metrics(ref, f, box), O.reference_metrics(ref, f, box)
    for k in mo:
        assert m[k] == pytest.approx(mo[k], rel=1e-12, abs=1e-14), k
    sp = 0.5
    assert S.sigma_j(f, sp, box) == pytest.approx(O.common_sigma_j(f, sp, box), rel=1e-12)
    Bv = np.stack([f[k] for k in ("Bx", "By", "Bz")])[:, box[0], box[1], box[2]]
    Jv = np.stack([f[k] for k in ("Jx", "Jy", "Jz")])[:, box[0], box[1], box[2]]
    assert S.roi_cwsin(Bv, Jv) == pytest.approx(O.box_cwsin(Bv, Jv), rel=1e-12)
    Rv = np.stack([ref[k] for k in ("Bx", "By", "Bz")])
    Fv = np.stack([f[k] for k in ("Bx", "By", "Bz")])
    want = np.sum(Fv * Rv) / (np.sqrt(np.sum(Fv ** 2) * np.sum(Rv ** 2)) + 1e-12)
    assert S.vector_correlation(Fv, Rv) == pytest.approx(want, rel=1e-12)


def test_metrics_reject_bad_boxes_and_non_finite_fields():
    from nfgs_amd import scores as S
    _announce("scores errors")
    f = _fields((4, 5, 6), 5)
    with pytest.raises(ValueError):
        S.compute_reference_metrics(f, f, (slice(0, 4), slice(0, 5, 2), slice(0, 6)))
    with pytest.raises(ValueError):
        S.compute_reference_metrics(f, f, (slice(2, 2), slice(0, 5), slice(0, 6)))
    g = dict(f)
    g["Bx"] = f["Bx"].copy()
    g["Bx"][1, 2, 3] = np.nan
    with pytest.raises(RuntimeError):
        S.compute_reference_metrics(f, g, (slice(None),) * 3)
