"""Pins of the oracle's common scoring (FP64 curl/div, sigma_J, agreement
metrics) by closed forms and by the reference's own known-answer case for the
common curl and divergence (tests/test_controlled_evaluation.py:13-33 of the
reference: a polynomial field on an anisotropic grid)."""
import numpy as np
import pytest

from oracle import nfgs_oracle as O


def _polynomial_field():
    x = np.linspace(-1.0, 1.0, 9)
    y = np.linspace(-2.0, 2.0, 7)
    z = np.linspace(0.0, 3.0, 5)
    zz, yy, xx = np.meshgrid(z, y, x, indexing="ij")
    return {"Bx": -yy, "By": xx, "Bz": 1.0 + xx * yy}, (0.25, 4.0 / 6.0, 0.75), xx, yy


def test_reference_known_answer_curl_and_divergence():
    f, sp, xx, yy = _polynomial_field()
    j = O.fd2_curl(f, sp)
    np.testing.assert_allclose(j["Jx"], xx, atol=2e-14)
    np.testing.assert_allclose(j["Jy"], -yy, atol=2e-14)
    np.testing.assert_allclose(j["Jz"], 2.0, atol=2e-14)
    np.testing.assert_allclose(O.fd2_divergence(f, sp), 0.0, atol=2e-14)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_derivative_exact_on_quadratics_with_one_sided_ends(axis):
    shape = [6, 5, 7]
    t = np.arange(shape[axis], dtype=np.float64) * 0.3
    f = 2.0 - 1.5 * t + 0.7 * t * t
    d = -1.5 + 1.4 * t
    g = np.moveaxis(np.broadcast_to(f, shape[:axis] + shape[axis + 1:] + [shape[axis]]), -1, axis)
    want = np.moveaxis(np.broadcast_to(d, shape[:axis] + shape[axis + 1:] + [shape[axis]]), -1,
                       axis)
    np.testing.assert_allclose(O.fd2_derivative(g, 0.3, axis), want, atol=1e-13)


def test_derivative_short_axes():
    f = np.array([[1.0, 4.0]])
    np.testing.assert_allclose(O.fd2_derivative(f, 0.5, 1), [[6.0, 6.0]])
    np.testing.assert_allclose(O.fd2_derivative(f, 0.5, 0), [[0.0, 0.0]])


def test_agreement_metrics_of_identical_and_scaled_fields():
    rng = np.random.default_rng(1)
    F = {k: rng.standard_normal((4, 5, 6)) for k in ("Bx", "By", "Bz", "Jx", "Jy", "Jz")}
    box = (slice(1, 4), slice(0, 5), slice(2, 6))
    m = O.reference_metrics(F, F, box)
    assert m["C_vec"] == pytest.approx(1.0, abs=1e-12)
    assert m["C_CS"] == pytest.approx(1.0, abs=1e-10)
    assert m["En"] == 0.0 and m["Em"] == 0.0
    assert m["epsilon"] == pytest.approx(1.0, abs=1e-12)
    G = dict(F)
    for k in ("Bx", "By", "Bz"):
        G[k] = 2.0 * F[k]
    m2 = O.reference_metrics(F, G, box)
    assert m2["epsilon"] == pytest.approx(4.0, rel=1e-12)
    assert m2["C_vec"] == pytest.approx(1.0, abs=1e-12)
    assert m2["En"] == pytest.approx(1.0, rel=1e-12)  # |2B - B| / |B|
    assert m2["Em"] == pytest.approx(1.0, rel=1e-10)


def test_cwsin_closed_forms():
    rng = np.random.default_rng(2)
    B = rng.standard_normal((3, 3, 4, 5)) + 3.0
    # J parallel to B: no Lorentz force
    assert O.box_cwsin(B, 0.7 * B) == pytest.approx(0.0, abs=1e-14)
    # J perpendicular to B everywhere: |J x B| / |B| = |J|
    e = np.array([1.0, -1.0, 0.0])[:, None, None, None] * np.ones_like(B)
    J = np.cross(B, e, axis=0)
    assert O.box_cwsin(B, J) == pytest.approx(1.0, rel=1e-12)
