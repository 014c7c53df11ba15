"""Host side of the controlled evaluation (nfgs_amd.controlled) and the
oracle's td_metrics.

The NPZ contract checks follow nfgs/controlled.py:25-52 of the reference and
its tests/test_controlled_evaluation.py:42-50 (coordinate and unit
mismatches).  The oracle's outward_errors (the host arithmetic of the
reference's cci_errors, controlled.py:55-68) is pinned by closed forms: a
zero target scores None, a candidate equal to the target scores 0 and one of
1.5 times the target scores 0.5 at every pixel and in L2.  td_metrics (the metrics of
recompute_td_referee_diagnostics.py:344-400) is pinned by its divergence term
on fields of known divergence, where the FD2 divergence is exact, and by the
identity case.
"""
import numpy as np
import pytest

from nfgs_amd.controlled import load, match
from oracle import nfgs_oracle as O


def _bundle(root):
    x = np.linspace(-1, 1, 9)
    y = np.linspace(-2, 2, 7)
    z = np.linspace(0, 3, 5)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    field = {"Bx": -yy, "By": xx, "Bz": 1 + xx * yy, "x": x, "y": y, "z": z,
             "field_unit": "benchmark", "length_unit": "normalized"}
    np.savez(root / "field.npz", **field)
    boundary = {k: field[k][0] for k in ("Bx", "By", "Bz")}
    boundary.update({k: field[k] for k in ("x", "y", "field_unit", "length_unit")})
    np.savez(root / "boundary.npz", **boundary)
    return field


def test_load_reads_the_contract(tmp_path):
    f = _bundle(tmp_path)
    a = load(tmp_path / "field.npz")
    assert a["field_unit"] == "benchmark" and a["length_unit"] == "normalized"
    np.testing.assert_array_equal(a["Bz"], f["Bz"])
    b = load(tmp_path / "boundary.npz", True)
    assert b["Bz"].shape == (7, 9) and "z" not in b
    match(a, b, True)


def test_coordinate_and_unit_mismatch(tmp_path):
    _bundle(tmp_path)
    a, b = load(tmp_path / "field.npz"), load(tmp_path / "field.npz")
    b["x"] = b["x"] + 1
    with pytest.raises(ValueError, match="Coordinate mismatch"):
        match(a, b)
    b = load(tmp_path / "field.npz")
    b["field_unit"] = "T"
    with pytest.raises(ValueError, match="Unit mismatch"):
        match(a, b)


def test_load_rejects_broken_files(tmp_path):
    f = _bundle(tmp_path)
    g = dict(f)
    del g["length_unit"]
    np.savez(tmp_path / "nounit.npz", **g)
    with pytest.raises(ValueError, match="missing coordinate/unit contract"):
        load(tmp_path / "nounit.npz")
    g = dict(f)
    g["x"] = np.array([0.0, 1.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0])
    np.savez(tmp_path / "nonuniform.npz", **g)
    with pytest.raises(ValueError, match="nonuniform x"):
        load(tmp_path / "nonuniform.npz")
    g = dict(f)
    g["Bx"] = g["Bx"][:, :, :-1]
    np.savez(tmp_path / "shape.npz", **g)
    with pytest.raises(ValueError, match="invalid Bx shape"):
        load(tmp_path / "shape.npz")
    g = dict(f)
    g["By"] = g["By"].copy()
    g["By"][1, 2, 3] = np.nan
    np.savez(tmp_path / "nan.npz", **g)
    with pytest.raises(ValueError, match="nonfinite"):
        load(tmp_path / "nan.npz")


def test_outward_errors_zero_target_and_closed_forms():
    cci_errors = O.outward_errors
    res = cci_errors(np.zeros((3, 7, 9)), np.ones((3, 7, 9)), (slice(None),) * 3, [1e-8, .01, .05])
    assert res["E_rel_Bgt_L2"] is None and res["E_rel_Bgt"] is None
    rng = np.random.default_rng(0)
    b = rng.standard_normal((3, 7, 9))
    box = (slice(None), slice(1, 6), slice(2, 8))
    same = cci_errors(b, b, box, [1e-8])
    assert same["E_rel_Bgt"] == 0.0 and same["E_rel_Bgt_L2"] == 0.0
    assert same["CCI_pixels_total"] == 30
    # a candidate of 1.5 b is 50% off at every pixel and in L2
    half = cci_errors(b, 1.5 * b, box, [1e-8, 0.5])
    assert half["E_rel_Bgt"] == pytest.approx(0.5, rel=1e-14)
    assert half["E_rel_Bgt_L2"] == pytest.approx(0.5, rel=1e-14)
    mag = np.linalg.norm(b[:, 1:6, 2:8], axis=0)
    kept = half["CCI_cutoff_results"]["0.5"]
    assert kept["retained_pixels"] == int(np.sum(mag > 0.5 * mag.max()))


def test_td_metrics_divergence_term_and_identity():
    x = np.linspace(0.0, 2.0, 9)
    y = np.linspace(0.0, 1.5, 7)
    z = np.linspace(0.0, 1.0, 5)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    h = (x[1] - x[0], y[1] - y[0], z[1] - z[0])
    # B = (x, 2y, -z) + (3, 0, 0): div B = 1 + 2 - 1 = 2, exact for FD2
    F = {"Bx": xx + 3.0, "By": 2.0 * yy, "Bz": -zz}
    div = O.fd2_divergence(F, h)
    np.testing.assert_allclose(div, 2.0, atol=1e-12)
    J = O.fd2_curl(F, h)
    box = (slice(1, 4), slice(0, 7), slice(2, 9))
    spacing = 3.0 / sum(1.0 / v for v in h)
    m = O.td_metrics(F, F, J, div, box, spacing)
    bn = np.sqrt((xx + 3.0) ** 2 + (2.0 * yy) ** 2 + zz ** 2)[box]
    assert m["divB_normalized"] == pytest.approx(2.0 / (6.0 * bn.mean() / spacing), rel=1e-12)
    assert m["C_vec"] == pytest.approx(1.0, abs=1e-12)
    assert m["Eprime_n"] == pytest.approx(1.0, abs=1e-12)
    assert m["Eprime_m"] == pytest.approx(1.0, abs=1e-12)
    assert m["epsilon"] == pytest.approx(1.0, abs=1e-12)
    # curl-free field: no current, sigma_J = 0 / floor
    assert m["sigma_J"] == pytest.approx(0.0, abs=1e-12)
