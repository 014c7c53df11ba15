"""Pins of the learned start's CPU oracle (oracle/fno_oracle.py): the
truncated DCT-II against scipy, the spectral layer's closed form, GELU, the
loss, Adam's first step, and every analytic gradient against central finite
differences in FP64."""
import numpy as np
import pytest
import scipy.fft
from scipy.stats import norm

from oracle import fno_oracle as fo


@pytest.mark.parametrize("n,m", [(7, 3), (16, 16), (130, 16)])
def test_truncated_dct2_is_scipys_orthonormal_dct(n, m):
    rng = np.random.default_rng(n)
    v = rng.normal(size=n)
    B = fo.dct2_basis(n, m)
    np.testing.assert_allclose(B.T @ v, scipy.fft.dct(v, type=2, norm="ortho")[:m], atol=1e-12)
    c = rng.normal(size=m)
    full = np.zeros(n)
    full[:m] = c
    np.testing.assert_allclose(B @ c, scipy.fft.idct(full, type=2, norm="ortho"), atol=1e-12)
    np.testing.assert_allclose(B.T @ B, np.eye(m), atol=1e-12)


def test_spectral_layer_with_all_modes_and_one_mix_is_pointwise():
    """With m = n on every axis the bases are orthogonal, so E(R P v) = R v
    when R is the same matrix for every mode."""
    rng = np.random.default_rng(1)
    c = fo.WIDTH
    v = rng.normal(size=(c, 3, 3, 3))
    r0 = rng.normal(size=(c, c))
    bases = tuple(fo.dct2_basis(3, 3) for _ in range(3))
    r = np.broadcast_to(r0, (27, c, c))
    _, s = fo._spectral(v, r, bases)
    out = np.einsum("czyk,xk->czyx", s, bases[2])
    np.testing.assert_allclose(out, np.einsum("oi,izyx->ozyx", r0, v), atol=1e-10)


def test_gelu_is_x_times_the_normal_cdf_and_its_slope():
    x = np.linspace(-6, 6, 241)
    np.testing.assert_allclose(fo.gelu(x), x * norm.cdf(x), atol=1e-14)
    h = 1e-6
    np.testing.assert_allclose(fo.gelu_grad(x), (fo.gelu(x + h) - fo.gelu(x - h)) / (2 * h),
                               atol=1e-8)


def test_loss_closed_form():
    out = np.zeros((3, 5, 4, 6))
    tgt = np.zeros_like(out)
    tgt[0, 2, 1, 1] = 3.0
    tgt[2, 0, 3, 5] = -2.0
    val, g = fo.loss(out, tgt)
    assert val == pytest.approx(13.0 / out.size, rel=1e-14)
    assert g[0, 2, 1, 1] == pytest.approx(-6.0 / out.size)
    assert g[2, 0, 3, 5] == pytest.approx(4.0 / out.size)
    assert np.count_nonzero(g) == 2


def test_adam_first_step_moves_by_lr_times_sign():
    rng = np.random.default_rng(2)
    w = rng.normal(size=50)
    g = rng.normal(size=50)
    new, m, v = fo.adam(w, np.zeros(50), np.zeros(50), g, 1, lr=1e-3, eps=0.0)
    np.testing.assert_allclose(new, w - 1e-3 * np.sign(g), atol=1e-15)


def test_bf16_round_is_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.1415927], np.float32)
    r = fo.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie, even mantissa kept
    assert r[2] == np.float32(1.0 + 4 * 2 ** -8)  # tie, rounds to even
    assert r[3] == 1.0
    assert abs(r[4] - np.float32(-3.140625)) == 0


def _tiny(seed=3, m=2, shape=(4, 5, 6)):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    params = {}
    for name, sh in fo.param_shapes(m):
        scale = 0.6 / np.sqrt(sh[-1]) if not name.endswith(".b") else 0.2
        params[name] = rng.normal(size=sh) * scale
    feat = rng.normal(size=(fo.FEATURES, nz, ny, nx))
    tgt = rng.normal(size=(fo.OUTPUTS, nz, ny, nx))
    bases = (fo.dct2_basis(nz, m), fo.dct2_basis(ny, m), fo.dct2_basis(nx, m))
    return params, feat, tgt, bases


def _loss_of(params, feat, tgt, bases):
    out, _ = fo.forward(params, feat, bases)
    return fo.loss(out, tgt)[0]


def test_every_gradient_matches_central_differences():
    params, feat, tgt, bases = _tiny()
    out, cache = fo.forward(params, feat, bases)
    _, gout = fo.loss(out, tgt)
    grads = fo.backward(params, cache, gout, bases)
    rng = np.random.default_rng(4)
    h = 1e-6
    for name, _ in fo.param_shapes(2):
        p = params[name]
        for _ in range(4):
            idx = tuple(int(rng.integers(0, s)) for s in p.shape)
            old = p[idx]
            p[idx] = old + h
            lp = _loss_of(params, feat, tgt, bases)
            p[idx] = old - h
            lm = _loss_of(params, feat, tgt, bases)
            p[idx] = old
            fd = (lp - lm) / (2 * h)
            assert grads[name][idx] == pytest.approx(fd, rel=2e-5, abs=1e-9), (name, idx)
    for _ in range(6):
        idx = tuple(int(rng.integers(0, s)) for s in feat.shape)
        old = feat[idx]
        feat[idx] = old + h
        lp = _loss_of(params, feat, tgt, bases)
        feat[idx] = old - h
        lm = _loss_of(params, feat, tgt, bases)
        feat[idx] = old
        assert grads["features"][idx] == pytest.approx((lp - lm) / (2 * h), rel=2e-5, abs=1e-9)


def test_a_planted_transpose_in_the_mix_gradient_fails_the_difference_check():
    """The finite-difference pin is sharp: the mix gradient with (o, i)
    swapped disagrees with it."""
    params, feat, tgt, bases = _tiny(seed=5)
    out, cache = fo.forward(params, feat, bases)
    _, gout = fo.loss(out, tgt)
    g = fo.backward(params, cache, gout, bases)["layer1.r"]
    wrong = np.swapaxes(g, 1, 2)
    h = 1e-6
    p = params["layer1.r"]
    idx = (3, 1, 7)
    old = p[idx]
    p[idx] = old + h
    lp = _loss_of(params, feat, tgt, bases)
    p[idx] = old - h
    lm = _loss_of(params, feat, tgt, bases)
    p[idx] = old
    fd = (lp - lm) / (2 * h)
    assert g[idx] == pytest.approx(fd, rel=2e-5)
    assert wrong[idx] != pytest.approx(fd, rel=1e-2)
