"""The oracle's outward Gaussian field B^> on z = 0.

gaussian_above_bottom_field restates nfgs_solver.py:4921-4943: the horizontal
part is -grad_h of dPhi/dz = idct2(k Phi^>_hat) with the spectral gradient of
:4385-4406, the vertical part is Bz_above from the split of :4810-4841.  The
reference's tests hold no vector for it, so these pins come from what the
mathematics fixes, on boundaries built analytically from endpoint-Neumann
cosine modes phi = A cos(a x) cos(b y) of the grid:

* a potential field of sources above the bottom (phi ~ exp(+k z)) is all
  outward: B^> returns the observed Bx, By and Bz exactly;
* a potential field of sources below (phi ~ exp(-k z)) has no outward part:
  B^> = 0;
* a sum of the two gives the first part alone.

k = sqrt(alpha_x^2 + alpha_y^2) and the derivative of cos(a x) is
-alpha sin(a x) with alpha = 2 sin(a dx / 2) / dx (the discrete-consistent
factor), so a wrong sign, a swapped axis, the below potential in place of the
above one or a dropped k fails these checks.
"""
import numpy as np
import pytest

from oracle import nfgs_oracle as O

NX, NY, NZ, DX, DY = 17, 13, 5, 1.0, 0.7


def _modes(seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(6):
        m = int(rng.integers(0, NX - 1))
        n = int(rng.integers(0, NY - 1))
        if m == 0 and n == 0:
            n = 1
        out.append((m, n, float(rng.standard_normal())))
    return out


def _potential_boundary(modes, sign):
    """(Bx, By, Bz) on z = 0 of B = -grad phi with phi = sum A cos cos exp(-sign k z):
    sign = +1 decays upward (sources below), sign = -1 decays downward."""
    x = np.arange(NX) * DX
    y = np.arange(NY) * DY
    bx = np.zeros((NY, NX))
    by = np.zeros((NY, NX))
    bz = np.zeros((NY, NX))
    for m, n, amp in modes:
        a = np.pi * m / ((NX - 1) * DX)
        b = np.pi * n / ((NY - 1) * DY)
        am = 2.0 * np.sin(np.pi * m / (2.0 * (NX - 1))) / DX
        an = 2.0 * np.sin(np.pi * n / (2.0 * (NY - 1))) / DY
        k = np.hypot(am, an)
        cx, sx = np.cos(a * x)[None, :], np.sin(a * x)[None, :]
        cy, sy = np.cos(b * y)[:, None], np.sin(b * y)[:, None]
        bx += amp * am * sx * cy
        by += amp * an * cx * sy
        bz += sign * amp * k * cx * cy
    return bx, by, bz


def _oracle(bx, by, bz):
    o = O.OracleNFGS(NX, NY, NZ, DX, DY, 1.0, dtype=np.float64, pad_calculation_domain=False,
                     source_zero_mode_policy="project", d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(bx, by, bz)
    return o


def test_sources_above_are_all_outward():
    bx, by, bz = _potential_boundary(_modes(1), -1.0)
    o = _oracle(bx, by, bz)
    assert (o.Nx, o.Ny) == (NX, NY)
    f = o.gaussian_above_bottom_field()
    assert f.shape == (3, NY, NX)
    scale = np.max(np.abs(np.stack([bx, by, bz])))
    np.testing.assert_allclose(f[0], bx, atol=1e-12 * scale)
    np.testing.assert_allclose(f[1], by, atol=1e-12 * scale)
    np.testing.assert_allclose(f[2], bz, atol=1e-12 * scale)


def test_sources_below_have_no_outward_part():
    bx, by, bz = _potential_boundary(_modes(2), 1.0)
    o = _oracle(bx, by, bz)
    f = o.gaussian_above_bottom_field()
    scale = np.max(np.abs(np.stack([bx, by, bz])))
    assert np.max(np.abs(f)) < 1e-12 * scale


@pytest.mark.parametrize("seed", [3, 4])
def test_mixed_sources_separate(seed):
    up = _potential_boundary(_modes(seed), -1.0)
    down = _potential_boundary(_modes(seed + 100), 1.0)
    o = _oracle(*(u + d for u, d in zip(up, down)))
    f = o.gaussian_above_bottom_field()
    scale = np.max(np.abs(np.stack(up)))
    for c in range(3):
        np.testing.assert_allclose(f[c], up[c], atol=1e-12 * scale)
    # B_z splits into the outward part and the below part of the background
    np.testing.assert_allclose(o.Bz_above + o.B_background[2, 0] - o.Bz_open_mean, o.Bz_obs,
                               atol=1e-12 * scale)
