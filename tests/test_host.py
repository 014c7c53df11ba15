"""Host-side logic of the drop-in (no GPU): grid sizing, the sigma window,
option validation, loud failure without a device, the input generators."""
import numpy as np
import pytest
import torch

from nfgs_amd import grid as G
from nfgs_amd.boundary import dipole_boundary, lfff_horizontal, multipole_boundary
from nfgs_amd.solver import NFGSSolverGPU, _Options
from oracle import nfgs_oracle as O


@pytest.mark.parametrize("n", [3, 5, 9, 12, 17, 33, 64, 65, 66, 100, 129, 150, 250, 256, 350, 512])
def test_grid_sizes_agree_with_oracle(n):
    assert G.mudpack_axis_size(n) == O.mudpack_axis_size(n)
    for mode in ("mudpack", "mudpack_legacy", "bottom_padding"):
        for pad in (True, False):
            try:
                ref = O.calculation_grid(n, n + 1, max(3, n // 2), pad, mode)
            except ValueError:
                with pytest.raises(ValueError):
                    G.calculation_grid(n, n + 1, max(3, n // 2), pad, mode)
                continue
            assert G.calculation_grid(n, n + 1, max(3, n // 2), pad, mode) == ref


@pytest.mark.parametrize("obs,calc", [((64, 64, 64), (66, 66, 65)), ((256, 256, 256), (258, 258, 257)),
                                      ((512, 512, 256), (514, 514, 257)), ((128, 128, 128), (130, 130, 129))])
def test_config_grids(obs, calc):
    nx, ny, nz, _ = G.calculation_grid(*obs, True, "mudpack")
    assert (nx, ny, nz) == calc


@pytest.mark.parametrize("obs", [(64, 64, 64), (12, 10, 9), (13, 11, 10), (256, 256, 256)])
def test_sigma_window_matches_oracle(obs):
    nxo, nyo, nzo = obs
    o = O.OracleNFGS(nxo, nyo, nzo)
    (sy0, sy1, sx0, sx1), (dy0, dy1, dx0, dx1) = G.centred_embedding(nyo, nxo, o.Ny, o.Nx)
    o.fov = (dy0, dy1, dx0, dx1)
    o.observed_zyx = (min(o.Nz, nzo), nyo, nxo)
    sz, sy, sx = o.window(o.Nz - 1)
    w = G.sigma_window(o.Nz - 1, o.Ny, o.Nx, o.observed_zyx, (dy0, dx0))
    r = lambda sl, n: range(*sl.indices(n))
    assert list(range(w[0], w[1])) == list(r(sz, o.Nz - 1))
    assert list(range(w[2], w[3])) == list(r(sy, o.Ny))
    assert list(range(w[4], w[5])) == list(r(sx, o.Nx))


def test_option_validation_mirrors_reference():
    with pytest.raises(ValueError, match="poisson_z_order"):
        _Options(dict(poisson_z_order=4))
    with pytest.raises(ValueError, match="mean_bz_policy"):
        _Options(dict(mean_bz_policy="nope"))
    with pytest.raises(ValueError, match="mutually exclusive"):
        _Options(dict(bottom_divergence_projection_weight=0.5, bottom_divergence_robin_weight=0.5))
    with pytest.raises(ValueError, match="alpha_clip_scale"):
        _Options(dict(alpha_clip_scale=2.0, alpha_limit_factor=1.0))
    with pytest.raises(NotImplementedError):
        _Options(dict(vertical_current_operator="collocated_centered"))
    o = _Options(dict(bottom_divergence_robin_weight=0.5, alpha_clip_scale=1.0))
    assert o.rho == 0.5 and o.alpha_limit_factor == 1.0


def test_coupled_closure_options():
    """Both per-iteration closures are accepted; the spline's relaxation
    defaults to 0.2 (nfgs_solver.py:4143-4147); the combinations the
    reference rejects, or this build does not run, raise before any device
    work."""
    for policy in ("coupled_k0", "constrained_spline_harmonic"):
        o = _Options(dict(horizontal_zero_mode_policy=policy))
        assert o.horizontal_zero_mode_policy == policy
        with pytest.raises(ValueError, match="relax_scalars"):
            _Options(dict(horizontal_zero_mode_policy=policy, relax_scalars=False))
        with pytest.raises(NotImplementedError):
            _Options(dict(horizontal_zero_mode_policy=policy, record_iteration_diagnostics=True))
        with pytest.raises(NotImplementedError):
            _Options(dict(horizontal_zero_mode_policy=policy,
                          k0_current_target_placement="face"))
    assert _Options(dict(horizontal_zero_mode_policy="constrained_spline_harmonic")
                    ).k0_update_mu == 0.2
    assert _Options(dict(horizontal_zero_mode_policy="coupled_k0")).k0_update_mu is None
    with pytest.raises(ValueError, match="k0_spline_basis_size"):
        NFGSSolverGPU(12, 10, 5, 1, 1, 1, horizontal_zero_mode_policy="constrained_spline_harmonic",
                      k0_spline_basis_size=12, pad_calculation_domain=False)
    with pytest.raises(ValueError, match="Nz >= 4"):
        NFGSSolverGPU(12, 10, 3, 1, 1, 1, horizontal_zero_mode_policy="coupled_k0",
                      pad_calculation_domain=False)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU failure mode")
def test_solver_fails_loudly_without_gpu():
    with pytest.raises(RuntimeError, match="GPU"):
        NFGSSolverGPU(12, 10, 9, 1, 1, 1)


def test_lfff_generator_is_force_free_and_divergence_consistent():
    bx, by, bz, alpha = multipole_boundary(64, seed=3, sigma=(3.0, 6.0))
    ny, nx = bz.shape
    kx = 2 * np.pi * np.fft.fftfreq(nx)
    ky = 2 * np.pi * np.fft.fftfreq(ny)
    KY, KX = np.meshgrid(ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2
    # Nyquist rows/columns are excluded: a real field cannot carry i*k there
    nyq = (np.abs(KX) >= np.pi - 1e-12) | (np.abs(KY) >= np.pi - 1e-12)
    ok = (k2 > alpha ** 2) & ~nyq
    bzh, bxh, byh = np.fft.fft2(bz), np.fft.fft2(bx), np.fft.fft2(by)
    curl = 1j * KX * byh - 1j * KY * bxh       # z of curl_h B_h
    np.testing.assert_allclose(curl[ok], alpha * bzh[ok], atol=1e-9 * np.abs(bzh).max())
    div = 1j * KX * bxh + 1j * KY * byh       # = l Bz for the decaying mode
    l = np.sqrt(np.where(ok, k2 - alpha ** 2, 0))
    np.testing.assert_allclose(div[ok], l[ok] * bzh[ok], atol=1e-9 * np.abs(bzh).max())


def test_multipole_flux_balance_and_dipole_shape():
    bx, by, bz, _ = multipole_boundary(256, seed=0)
    # the poles balance sum(a_i sigma_i^2) exactly; the discrete sum differs
    # only by the tails the domain edge cuts off
    assert abs(bz.sum()) < 1e-3 * np.abs(bz).sum()
    bx, by, bz = dipole_boundary(64, seed=1)
    assert bz.shape == (64, 64) and bz.max() > 0.9 and bz.min() < -0.9
