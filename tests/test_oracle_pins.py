"""Pins of the oracle that do not depend on the reference: library routines,
closed forms, dense solves, and the reference's own known-answer tests."""
import numpy as np
import pytest
import scipy.fft

from oracle import nfgs_oracle as O


@pytest.mark.parametrize("n", [3, 4, 9, 14, 66])
def test_dct1_dst1_match_scipy_and_closed_form(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, n))
    np.testing.assert_allclose(O.dct1(x, -1), scipy.fft.dct(x, type=1, axis=-1), atol=1e-12)
    np.testing.assert_allclose(O.dst1(x, -1), scipy.fft.dst(x, type=1, axis=-1), atol=1e-12)
    j = np.arange(n)
    C = 2 * np.cos(np.pi * np.outer(j, j) / (n - 1))
    C[:, 0] = 1
    C[:, -1] = (-1.0) ** j
    S = 2 * np.sin(np.pi * np.outer(j + 1, j + 1) / (n + 1))
    np.testing.assert_allclose(O.dct1(x, -1), x @ C.T, atol=1e-11)
    np.testing.assert_allclose(O.dst1(x, -1), x @ S.T, atol=1e-11)
    # involutions
    np.testing.assert_allclose(O.idct1(O.dct1(x, -1), -1), x, atol=1e-12)
    np.testing.assert_allclose(O.idst1(O.dst1(x, -1), -1), x, atol=1e-12)
    # transforms along a non-last axis agree with the last-axis form
    np.testing.assert_allclose(O.dct1(x.T, 0), O.dct1(x, -1).T, atol=1e-12)


def test_lambda_tables_are_eigenvalues_of_fd_laplacian():
    s = O.OracleNFGS(10, 8, 5, 0.7, 1.3, 1.0)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((s.Ny, s.Nx))
    lhs = O.dct2(O.lap_xy(f, s.dx, s.dy))
    np.testing.assert_allclose(lhs, s.lam_neumann() * O.dct2(f), atol=1e-10)
    g = np.zeros_like(f)
    g[1:-1, 1:-1] = rng.standard_normal((s.Ny - 2, s.Nx - 2))
    lhs = O.dst2(O.lap_xy(g, s.dx, s.dy)[1:-1, 1:-1])
    np.testing.assert_allclose(lhs, s.lam_dirichlet() * O.dst2(g[1:-1, 1:-1]), atol=1e-10)
    # alpha^2 = -lambda per axis
    ax, ay = s.alpha_x(), s.alpha_y()
    np.testing.assert_allclose(-(ay[:, None] ** 2 + ax[None, :] ** 2), s.lam_neumann(), atol=1e-12)


def test_fd_stencils_exact_on_quadratics():
    z, y, x = np.meshgrid(np.arange(6.0) * 0.5, np.arange(7.0) * 0.25, np.arange(8.0) * 0.4,
                          indexing="ij")
    f = 1 + 2 * x + 3 * y - z + x * x - 0.5 * y * y + 0.7 * z * z + x * z - 2 * y * z
    np.testing.assert_allclose(O.d_dx(f, 0.4), 2 + 2 * x + z, atol=1e-11)
    np.testing.assert_allclose(O.d_dy(f, 0.25), 3 - y - 2 * z, atol=1e-11)
    np.testing.assert_allclose(O.d_dz(f, 0.5), -1 + 1.4 * z + x - 2 * y, atol=1e-11)
    np.testing.assert_allclose(O.d2_dxdz(f, 0.4, 0.5)[1:-1, :, 1:-1], 1.0, atol=1e-11)
    np.testing.assert_allclose(O.d2_dydz(f, 0.25, 0.5)[1:-1, 1:-1, :], -2.0, atol=1e-11)
    np.testing.assert_allclose(O.lap_xy(f, 0.4, 0.25)[:, 1:-1, 1:-1], 2 - 1.0, atol=1e-10)


def _dense_dirichlet(rhs, u0, ut, lam, dz):
    n = rhs.shape[0]
    m = n - 2
    A = np.zeros((m, m))
    b = rhs[1:-1].copy()
    for i in range(m):
        A[i, i] = lam - 2 / dz ** 2
        if i > 0:
            A[i, i - 1] = 1 / dz ** 2
        if i < m - 1:
            A[i, i + 1] = 1 / dz ** 2
    b[0] -= u0 / dz ** 2
    b[-1] -= ut / dz ** 2
    return np.concatenate([[u0], np.linalg.solve(A, b), [ut]])


@pytest.mark.parametrize("nz", [3, 4, 11])
def test_thomas_dirichlet_matches_dense_solve(nz):
    rng = np.random.default_rng(nz)
    lam = -np.abs(rng.standard_normal((2, 3))) - 0.1
    rhs = rng.standard_normal((nz, 2, 3))
    u0, ut = rng.standard_normal((2, 3)), rng.standard_normal((2, 3))
    u = O.thomas_dirichlet(rhs, u0, ut, lam, 0.7)
    for j in range(2):
        for i in range(3):
            ref = _dense_dirichlet(rhs[:, j, i], u0[j, i], ut[j, i], lam[j, i], 0.7)
            np.testing.assert_allclose(u[:, j, i], ref, atol=1e-10)


def _dense_robin(s, b, q, g, lam, k, dz, rho):
    """Assemble the full (Nz-1) system of the modal Robin solve for one mode:
    PDE rows 1..N-1 on planes, plane N-1's row replaced by the second-order
    top Robin condition, and the first row blended with the bottom derivative
    condition (-3u0+4u1-u2)/(2dz) = q with weight rho (scaled by 1/dz)."""
    n = s.shape[0]
    m = n - 1  # unknowns u1..u_{N-1}
    A = np.zeros((m, m))  # unknown planes 1..N map to indices 0..m-1
    r = np.zeros(m)
    h2 = dz * dz
    for p in range(1, n - 1):  # PDE rows at planes 1..N-2
        i = p - 1
        w_pde = (1 - rho) if p == 1 else 1.0
        A[i, i] += w_pde * (lam - 2 / h2)
        if i + 1 < m:
            A[i, i + 1] += w_pde / h2
        if i - 1 >= 0:
            A[i, i - 1] += w_pde / h2
        r[i] += w_pde * s[p]
        if p == 1:
            r[i] -= w_pde * b / h2
            # -(rho/dz) * ((-3 u0 + 4 u1 - u2)/(2 dz) - q) = 0
            A[0, 0] += -rho / dz * 4 / (2 * dz)
            if m > 1:
                A[0, 1] += -rho / dz * (-1) / (2 * dz)
            # known u0 = b and the target q move to the right-hand side
            r[0] += -rho * q / dz - 3.0 * rho * b / (2.0 * dz * dz)
    # top: (3u_N - 4u_{N-1} + u_{N-2})/(2dz) + k u_N = g  with N = n-1, u_N = unknown m-1
    N = m - 1
    A[N, :] = 0
    A[N, N] = 3 / (2 * dz) + k
    A[N, N - 1] += -4 / (2 * dz)
    if N - 2 >= 0:
        A[N, N - 2] += 1 / (2 * dz)
    rr = g
    if N - 2 < 0:
        rr -= b / (2 * dz)
    r[N] = rr
    return np.concatenate([[b], np.linalg.solve(A, r)])


@pytest.mark.parametrize("nz", [4, 5, 12])
@pytest.mark.parametrize("rho", [0.0, 0.4, 1.0])
def test_modal_robin_matches_dense_solve(nz, rho):
    rng = np.random.default_rng(nz + int(10 * rho))
    ny, nx, dz = 3, 4, 0.8
    lam = -np.abs(rng.standard_normal((ny, nx))) - 0.2
    lam[0, 0] = 0.0
    k = np.sqrt(-lam)
    s = rng.standard_normal((nz, ny, nx))
    b = rng.standard_normal((ny, nx))
    q = rng.standard_normal((ny, nx))
    g = rng.standard_normal((ny, nx))
    u = O.modal_robin_from_hat(s, b, lam, k, dz, rho, q, top_hat=g)
    assert np.all(u[:, 0, 0] == 0)
    for j in range(ny):
        for i in range(nx):
            if i == 0 and j == 0:
                continue
            ref = _dense_robin(s[:, j, i], b[j, i], q[j, i], g[j, i], lam[j, i], k[j, i], dz, rho)
            np.testing.assert_allclose(u[:, j, i], ref, rtol=1e-9, atol=1e-9)


def test_jz_linear_field_interior_curl_and_closed_edges():
    # tests/test_bottom_current_padding.py:11-19 of the reference, restated
    y, x = np.meshgrid(np.arange(8.0) * .7, np.arange(10.0) * .3, indexing="ij")
    jz = O.derive_bottom_jz_calc(2 * y, 5 * x, .3, .7, ijzbcor=1)
    np.testing.assert_allclose(jz[1:-1, 1:-1], 3.0)
    assert np.count_nonzero(jz[[0, -1], :]) == 0
    assert np.count_nonzero(jz[:, [0, -1]]) == 0


@pytest.mark.parametrize("mode", [2, 3, 4])
def test_jz_balance_preserves_signs_and_zero_cell_quadrature(mode):
    # tests/test_bottom_current_padding.py:22-35 of the reference, restated
    y, x = np.meshgrid(np.arange(8.0), np.arange(10.0), indexing="ij")
    bx, by = .2 * y ** 2, np.sin(x) * (1 + y)
    raw = O.derive_bottom_jz_calc(bx, by, 1., 1., ijzbcor=1)
    bal = O.derive_bottom_jz_calc(bx, by, 1., 1., ijzbcor=mode)
    assert np.array_equal(np.sign(raw), np.sign(bal))
    cell = .25 * (bal[1:-1, 1:-1] + bal[2:, 1:-1] + bal[1:-1, 2:] + bal[2:, 2:])
    assert abs(float(cell.sum())) < 1e-12


def test_legacy_helper_jz_zero_edges_of_helper_grid_only():
    # The helper grid is larger than the calculation grid; its edge nodes are
    # zeroed, but the returned calculation-grid edges are not forced to zero.
    rng = np.random.default_rng(3)
    bx, by = rng.standard_normal((14, 14)), rng.standard_normal((14, 14))
    jz = O.derive_bottom_jz_legacy(bx, by, 1.0, 1.0, 9, ijzbcor=1)
    g = O.fortran_grid(14, 14, 9)
    assert (g["ny"], g["nx"]) == (17, 17)
    assert np.count_nonzero(jz[:, 0]) > 0


@pytest.mark.parametrize("nobs,calc", [(64, 66), (256, 258), (512, 514), (80, 98), (150, 194),
                                       (250, 258), (350, 386), (128, 130), (96, 98), (320, 386)])
def test_calculation_grid_matches_archived_sizes(nobs, calc):
    cx, _, _, _ = O.calculation_grid(nobs, nobs, nobs, True, "mudpack")
    assert cx == calc


def test_gaussian_split_invariants():
    rng = np.random.default_rng(5)
    s = O.OracleNFGS(16, 14, 10)
    y, x = np.meshgrid(np.arange(14.0), np.arange(16.0), indexing="ij")
    bz = np.exp(-((x - 5) ** 2 + (y - 6) ** 2) / 8) - np.exp(-((x - 10) ** 2 + (y - 7) ** 2) / 8)
    bx = 0.1 * np.sin(0.3 * y) + 0.01 * rng.standard_normal(bz.shape)
    by = 0.05 * np.cos(0.2 * x)
    s.set_bottom_vector_boundary(bx, by, bz)
    # bottom Bz is reconstructed exactly by background plus above-field
    np.testing.assert_allclose(s.B_background[2, 0] + s.Bz_above, s.Bz_obs, atol=1e-12)
    # spectral divergence has an exactly zero constant mode
    assert abs(s.gaussian_diagnostics["D0_zero_mode"]) < 1e-14
    # the below field decays upward mode by mode: its DCT2 shrinks like exp(-k z)
    h0 = O.dct2(s.B_background[2, 1] - s.Bz_open_mean)
    h1 = O.dct2(s.B_background[2, 2] - s.Bz_open_mean)
    mask = np.abs(h0) > 1e-8
    np.testing.assert_allclose((h1[mask] / h0[mask]), np.exp(-s.k[mask] * s.dz), rtol=1e-8)


@pytest.mark.parametrize("n", [3, 9, 66, 130])
def test_dense_and_threaded_transform_arithmetic_match_the_fft_form(n):
    """The oracle's alternative arithmetic for the type-I transforms (dense
    closed-form matrices; threaded FFTs) computes the same transforms: equal
    to the mirrored-FFT form to FP64 roundoff and to scipy's type-1 routines."""
    import scipy.fft as sf
    rng = np.random.default_rng(n)
    x = rng.standard_normal((4, n, n + 1))
    ref = {(k, a): f(x, a) for k, f in (("c", O.dct1), ("s", O.dst1)) for a in (-1, -2)}
    try:
        for mode in ("dense", "threads"):
            if mode == "dense":
                O.set_transform("dense")
            else:
                O.set_transform("fft")
                O.set_fft_workers(4)
            for (k, a), want in ref.items():
                got = (O.dct1 if k == "c" else O.dst1)(x, a)
                scale = np.abs(want).max()
                assert np.abs(got - want).max() <= 1e-13 * scale, (mode, k, a)
                lib = (sf.dct if k == "c" else sf.dst)(x, type=1, axis=a)
                assert np.abs(got - lib).max() <= 1e-13 * scale, (mode, k, a)
    finally:
        O.set_transform("fft")
        O.set_fft_workers(None)
    with pytest.raises(ValueError):
        O.set_transform("fp32")
