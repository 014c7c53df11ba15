"""CPU oracle for the NFGS plain iterate — TEST INFRASTRUCTURE ONLY.

This module is a NumPy restatement of the reference solver's default NFGS
path (Gaussian-separated background, frozen-linear-harmonic closure,
staggered-face vertical current, direct poloidal update, gs_cfcf start,
sigma_best stopping).  It exists to check the HIP product path.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import it, and only as the checker or the timed CPU baseline; the
product package never imports it.

Every function cites the reference lines it follows (``nfgs_solver.py``).

Parity pins (what checks each part; see tests/test_oracle_*.py):
  * transforms           scipy.fft.dct/dst(type=1), dense closed-form sums,
                         involution C.C = 2(n-1) I and S.S = 2(n+1) I
  * lambda tables / FD   DCT2(lap_xy f) == lambda * DCT2(f); polynomial
                         exactness of the FD stencils
  * z solves             dense numpy.linalg.solve of the assembled systems
  * Jz derivation        the reference's own known-answer tests
                         (linear field curl 3.0; sign-preserving balance)
  * everything composite golden fixtures written by tests/golden/make_golden.py
                         (reference imported on CPU, FP64 and FP32)
  * project_field        round trips on representable states, the dropped
                         per-plane constant mode (test_oracle_handoff.py)
  * common scoring       the reference's known-answer polynomial field and
                         closed forms of the metrics (test_oracle_scores.py)
  * outward field        potential fields of sources above / below the
                         bottom plane (test_oracle_above_field.py)
  * td_metrics           the divergence term on fields of known divergence
                         (test_controlled_contract.py)
  * coupled closures     reference vectors (tests/golden/make_k0.py), the
                         profile solves' normal equations (test_oracle_k0.py)
  * soft_cfcf start      reference vectors (tests/golden/make_soft.py); the
                         exact rows, minimality of the modal form and the
                         eps = 0 closed form (test_oracle_soft.py)
No function is left "parity unpinned".
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

# ----------------------------------------------------------------------------
# Grid sizing (nfgs_solver.py:73-177, :771-819)
# ----------------------------------------------------------------------------


def mudpack_axis_size(n_obs: int) -> int:
    """nfgs_solver.py:73-83."""
    n_obs = int(n_obs)
    if n_obs < 2:
        return max(1, n_obs)
    n = float(n_obs)
    e2 = math.ceil(math.log((n - 1.0) / 2.0) / math.log(2.0)) + 1
    e3 = math.ceil(math.log((n - 1.0) / 3.0) / math.log(2.0)) + 1
    return min(int(2 * 2 ** (e2 - 1) + 1), int(3 * 2 ** (e3 - 1) + 1))


def even_xy_pad_size(n_calc: int, n_obs: int) -> int:
    """nfgs_solver.py:86-102."""
    if n_calc < n_obs:
        raise ValueError(f"calculation size {n_calc} is smaller than observed size {n_obs}")
    return n_calc + 1 if (n_calc - n_obs) % 2 == 1 else n_calc


def fortran_grid(nxo: int, nyo: int, nzo: int, symmetric_xy: bool = False):
    """Sizes and pads of nfgs_solver.py:105-177 (spacings are not needed)."""
    nx, ny, nz = mudpack_axis_size(nxo), mudpack_axis_size(nyo), mudpack_axis_size(nzo)
    if symmetric_xy:
        nx, ny = even_xy_pad_size(nx, nxo), even_xy_pad_size(ny, nyo)
    return dict(nx=nx, ny=ny, nz=nz, bfx=(nx - nxo) // 2, bfy=(ny - nyo) // 2, bfz=nz - nzo)


def calculation_grid(Nx, Ny, Nz, pad=True, mode="mudpack"):
    """Calculation grid and lateral pads, nfgs_solver.py:771-819."""
    mode = str(mode).lower().replace("-", "_")
    if mode not in ("mudpack", "mudpack_legacy", "bottom_padding"):
        raise ValueError(f"calculation_domain_mode invalid: {mode!r}")
    if not pad:
        return Nx, Ny, Nz, (0, 0)
    if mode in ("mudpack", "mudpack_legacy"):
        g = fortran_grid(Nx, Ny, Nz, symmetric_xy=(mode == "mudpack"))
        return g["nx"], g["ny"], g["nz"], (g["bfx"], g["bfy"])
    p = (Nx + Ny) // 6
    cx, cy = Nx + 2 * p, Ny + 2 * p
    cz = min(cx, cy)
    if cz < Nz:
        raise ValueError("bottom_padding calculation height cannot contain the input z domain")
    return cx, cy, cz, (p, p)


# ----------------------------------------------------------------------------
# Type-I transforms via mirrored real FFTs (nfgs_solver.py:1019-1081)
# ----------------------------------------------------------------------------


def _move(x, axis):
    return np.moveaxis(x, axis, -1)


# Threads of the real FFTs behind the type-I transforms: None keeps numpy's
# single-threaded rfft (the setting every parity test uses); a count n > 1
# switches to scipy.fft.rfft(..., workers=n), the same pocketfft algorithm on
# n threads, for the timed CPU baseline.
_FFT_WORKERS = None


def set_fft_workers(n):
    """Use n threads for the transforms' FFTs (None or 1: numpy, one thread)."""
    global _FFT_WORKERS
    _FFT_WORKERS = None if n is None or int(n) <= 1 else int(n)


def _rfft(y):
    if _FFT_WORKERS is None:
        return np.fft.rfft(y, axis=-1)
    import scipy.fft
    return scipy.fft.rfft(y, axis=-1, workers=_FFT_WORKERS)


# Arithmetic of the type-I transforms: "fft" (mirrored real FFTs, the
# reference's form and the default) or "dense" (products with the closed-form
# matrices through BLAS).  The two are the same transforms in independent
# FP64 roundings; the spread between whole solves run each way is the FP64
# floor that conditioning alone sets at a grid size.
_TRANSFORM = "fft"
_DENSE = {}


def set_transform(kind):
    """Select "fft" (default) or "dense" transform arithmetic."""
    global _TRANSFORM
    if kind not in ("fft", "dense"):
        raise ValueError("transform arithmetic must be 'fft' or 'dense'")
    _TRANSFORM = kind


def _dense_matrix(kind, n):
    key = (kind, n)
    if key not in _DENSE:
        j = np.arange(n)[None, :]
        k = np.arange(n)[:, None]
        if kind == "dct":
            m = 2.0 * np.cos(np.pi * j * k / (n - 1))
            m[:, 0] = 1.0
            m[:, -1] = (-1.0) ** np.arange(n)
        else:
            m = 2.0 * np.sin(np.pi * (j + 1) * (k + 1) / (n + 1))
        _DENSE[key] = m.T.copy()
    return _DENSE[key]


def _dense(kind, x, axis):
    x = np.asarray(x)
    xm = np.moveaxis(x, axis, -1)
    return np.moveaxis(xm @ _dense_matrix(kind, xm.shape[-1]), -1, axis).astype(x.dtype)


def dst1(x, axis):
    """Unnormalised DST-I, nfgs_solver.py:1019-1045."""
    if _TRANSFORM == "dense":
        return _dense("dst", x, axis)
    x = _move(np.asarray(x), axis)
    n = x.shape[-1]
    y = np.zeros(x.shape[:-1] + (2 * (n + 1),), dtype=x.dtype)
    y[..., 1:n + 1] = x
    y[..., n + 2:] = -x[..., ::-1]
    Y = _rfft(y)
    return np.moveaxis((-Y.imag[..., 1:n + 1]).astype(x.dtype), -1, axis)


def idst1(X, axis):
    """nfgs_solver.py:1047-1050."""
    n = np.asarray(X).shape[axis]
    return dst1(X, axis) / (2.0 * (n + 1))


def dct1(x, axis):
    """Unnormalised DCT-I, nfgs_solver.py:1052-1074."""
    if _TRANSFORM == "dense" and np.asarray(x).shape[axis] > 1:
        return _dense("dct", x, axis)
    x = _move(np.asarray(x), axis)
    n = x.shape[-1]
    if n == 1:
        return np.moveaxis(x.copy(), -1, axis)
    y = np.empty(x.shape[:-1] + (2 * (n - 1),), dtype=x.dtype)
    y[..., :n] = x
    y[..., n:] = x[..., n - 2:0:-1]
    Y = _rfft(y)
    return np.moveaxis(Y.real[..., :n].astype(x.dtype), -1, axis)


def idct1(X, axis):
    """nfgs_solver.py:1076-1081."""
    n = np.asarray(X).shape[axis]
    if n == 1:
        return np.array(X, copy=True)
    return dct1(X, axis) / (2.0 * (n - 1))


def dct2(v):
    """nfgs_solver.py:4354-4355."""
    return dct1(dct1(v, -1), -2)


def idct2(v):
    """nfgs_solver.py:4459-4460."""
    return idct1(idct1(v, -1), -2)


def dst2(v):
    return dst1(dst1(v, -1), -2)


def idst2(v):
    return idst1(idst1(v, -1), -2)


# ----------------------------------------------------------------------------
# Finite differences (nfgs_solver.py:1361-1465, :1711-1768)
# ----------------------------------------------------------------------------


def _d_axis(f, h, axis):
    """Second-order centred derivative with one-sided ends (fdo=2 branch)."""
    f = np.moveaxis(f, axis, -1)
    out = np.zeros_like(f)
    n = f.shape[-1]
    if n < 2:
        return np.moveaxis(out, -1, axis)
    if n >= 3:
        out[..., 1:-1] = (f[..., 2:] - f[..., :-2]) / (2.0 * h)
        out[..., 0] = (-3.0 * f[..., 0] + 4.0 * f[..., 1] - f[..., 2]) / (2.0 * h)
        out[..., -1] = (3.0 * f[..., -1] - 4.0 * f[..., -2] + f[..., -3]) / (2.0 * h)
    else:
        out[..., 0] = (f[..., 1] - f[..., 0]) / h
        out[..., -1] = (f[..., -1] - f[..., -2]) / h
    return np.moveaxis(out, -1, axis)


def d_dx(f, dx):
    """nfgs_solver.py:1361-1379 (also the 2-D twin :1428-1446)."""
    return _d_axis(f, dx, -1)


def d_dy(f, dy):
    """nfgs_solver.py:1381-1399 (also :1448-1465)."""
    return _d_axis(f, dy, -2)


def d_dz(f, dz):
    """nfgs_solver.py:1401-1419."""
    return _d_axis(f, dz, 0)


def lap_xy(f, dx, dy):
    """Mirror-Neumann 5-point lateral Laplacian, nfgs_solver.py:1741-1749."""
    lap = np.zeros_like(f)
    lap[..., 1:-1] += (f[..., 2:] - 2 * f[..., 1:-1] + f[..., :-2]) / dx ** 2
    lap[..., 1:-1, :] += (f[..., 2:, :] - 2 * f[..., 1:-1, :] + f[..., :-2, :]) / dy ** 2
    lap[..., 0, :] += (2 * f[..., 1, :] - 2 * f[..., 0, :]) / dy ** 2
    lap[..., -1, :] += (2 * f[..., -2, :] - 2 * f[..., -1, :]) / dy ** 2
    lap[..., 0] += (2 * f[..., 1] - 2 * f[..., 0]) / dx ** 2
    lap[..., -1] += (2 * f[..., -2] - 2 * f[..., -1]) / dx ** 2
    return lap


def d2_dxdz(f, dx, dz):
    """nfgs_solver.py:1729-1733."""
    d2 = np.zeros_like(f)
    if f.shape[0] < 3 or f.shape[2] < 3:
        return d2
    d2[1:-1, :, 1:-1] = (f[2:, :, 2:] - f[:-2, :, 2:] - f[2:, :, :-2] + f[:-2, :, :-2]) / (4.0 * dx * dz)
    return d2


def d2_dydz(f, dy, dz):
    """nfgs_solver.py:1735-1739."""
    d2 = np.zeros_like(f)
    if f.shape[0] < 3 or f.shape[1] < 3:
        return d2
    d2[1:-1, 1:-1, :] = (f[2:, 2:, :] - f[:-2, 2:, :] - f[2:, :-2, :] + f[:-2, :-2, :]) / (4.0 * dy * dz)
    return d2


def d2_dz2(f, dz):
    """Second z derivative, nfgs_solver.py:1717-1727: centred inside; at the
    bottom the four-point one-sided form (two-point below 4 planes); at the
    top the mirrored form."""
    d2 = np.zeros_like(f)
    dz2 = dz * dz
    if f.shape[0] >= 3:
        d2[1:-1] = (f[2:] - 2 * f[1:-1] + f[:-2]) / dz2
    if f.shape[0] >= 2:
        if f.shape[0] >= 4:
            d2[0] = (2.0 * f[0] - 5.0 * f[1] + 4.0 * f[2] - f[3]) / dz2
        else:
            d2[0] = (f[1] - f[0]) / dz2
        d2[-1] = (-2.0 * f[-1] + 2.0 * f[-2]) / dz2
    return d2


def reconstruct_pt(phi, az, bcz, b_pot, bz_obs, jz_bottom, dx, dy, dz):
    """Collocated PT state from the scalars with ccu_pure_current
    (nfgs_solver.py:3311-3341, :1711-1715, :1760-1768; boundary values
    :6069-6078, :3255-3257).  Returns (B, J, B_c)."""
    dphi = d_dz(phi, dz)
    bcx = -d_dx(dphi, dx) + d_dy(az, dy)
    bcy = -d_dy(dphi, dy) - d_dx(az, dx)
    daz = d_dz(az, dz)
    lphi = bcz + d2_dz2(phi, dz)
    jx = d_dx(daz, dx) + d_dy(lphi, dy)
    jy = d_dy(daz, dy) - d_dx(lphi, dx)
    jz = -lap_xy(az, dx, dy)
    Bc = np.stack([bcx, bcy, bcz])
    J = np.stack([jx, jy, jz])
    B = b_pot + Bc
    J[:, :, :, 0] = 0.0
    J[:, :, :, -1] = 0.0
    J[:, :, 0, :] = 0.0
    J[:, :, -1, :] = 0.0
    J[:, -1] = 0.0
    B[2, 0] = bz_obs
    J[2, 0] = jz_bottom
    return B, J, Bc


def pad_or_crop_center_2d(a2, ty, tx):
    """nfgs_solver.py:1593-1609; returns (out, (y0, y1, x0, x1))."""
    ny, nx = a2.shape
    out = np.zeros((ty, tx), dtype=a2.dtype)
    sy0, sx0 = max(0, (ny - ty) // 2), max(0, (nx - tx) // 2)
    sy1, sx1 = sy0 + min(ny, ty), sx0 + min(nx, tx)
    dy0, dx0 = max(0, (ty - ny) // 2), max(0, (tx - nx) // 2)
    dy1, dx1 = dy0 + (sy1 - sy0), dx0 + (sx1 - sx0)
    out[dy0:dy1, dx0:dx1] = a2[sy0:sy1, sx0:sx1]
    return out, (dy0, dy1, dx0, dx1)


# ----------------------------------------------------------------------------
# Bottom Jz (nfgs_solver.py:1498-1567 legacy helper; :4489-4546 dispatch)
# ----------------------------------------------------------------------------


def _cell_sum(a):
    return 0.25 * (a[1:-1, 2:].sum() + a[2:, 1:-1].sum() + a[1:-1, 1:-1].sum() + a[2:, 2:].sum())


def _balance(jz, mode, legacy):
    pmask = jz >= 0.0
    pos = np.where(jz >= 0.0, jz, 0.0).astype(jz.dtype)
    neg = np.where(jz < 0.0, jz, 0.0).astype(jz.dtype)
    tot, ps, ns = _cell_sum(jz), _cell_sum(pos), _cell_sum(neg)
    if mode == 2:
        tgt = ps - 0.5 * tot
    elif mode == 3:
        tgt = max(ps, abs(ns))
    elif mode == 4:
        tgt = min(ps, abs(ns))
    else:
        return jz
    if legacy:
        # masks are taken before either rescale (nfgs_solver.py:1556-1561)
        if abs(ps) > 0.0:
            jz = np.where(pmask, (tgt / ps) * jz, jz)
        if abs(ns) > 0.0:
            jz = np.where(~pmask, (tgt / abs(ns)) * jz, jz)
        return jz
    fp = tgt / ps if abs(ps) > 0 else 1.0
    fn = tgt / abs(ns) if abs(ns) > 0 else 1.0
    return np.where(jz >= 0, fp * jz, fn * jz)


def derive_bottom_jz_legacy(bx, by, dx, dy, nz, ijzbcor=3):
    """Legacy helper-grid Jz, nfgs_solver.py:1498-1567."""
    nyo, nxo = bx.shape
    g = fortran_grid(nxo, nyo, max(2, int(nz)))
    bxb = np.zeros((g["ny"], g["nx"]), dtype=bx.dtype)
    byb = np.zeros_like(bxb)
    y1, x1 = g["bfy"] + nyo, g["bfx"] + nxo
    bxb[g["bfy"]:y1, g["bfx"]:x1] = bx
    byb[g["bfy"]:y1, g["bfx"]:x1] = by
    jzb = np.zeros_like(bxb)
    if g["nx"] >= 3 and g["ny"] >= 3:
        jzb[1:-1, 1:-1] = ((byb[1:-1, 2:] - byb[1:-1, :-2]) / (2.0 * dx)
                           - (bxb[2:, 1:-1] - bxb[:-2, 1:-1]) / (2.0 * dy))
        if int(ijzbcor) >= 2:
            jzb = _balance(jzb, int(ijzbcor), legacy=True).astype(bx.dtype)
    jzb[0, :] = 0.0
    jzb[-1, :] = 0.0
    jzb[:, 0] = 0.0
    jzb[:, -1] = 0.0
    return jzb[g["bfy"]:y1, g["bfx"]:x1].copy()


def derive_bottom_jz_calc(bx, by, dx, dy, ijzbcor=3):
    """Unpadded-domain Jz on the calculation grid, nfgs_solver.py:4519-4546."""
    jz = np.zeros_like(bx)
    ny, nx = bx.shape
    if nx >= 3 and ny >= 3:
        jz[1:-1, 1:-1] = ((by[1:-1, 2:] - by[1:-1, :-2]) / (2 * dx)
                          - (bx[2:, 1:-1] - bx[:-2, 1:-1]) / (2 * dy))
        if int(ijzbcor) >= 2:
            jz = _balance(jz, int(ijzbcor), legacy=False).astype(bx.dtype)
    return jz


# ----------------------------------------------------------------------------
# Batched z solves (nfgs_solver.py:535-608, :5512-5625)
# ----------------------------------------------------------------------------


def thomas_dirichlet(rhs, u0, utop, lam, dz):
    """_thomas_solve_batched_lowmem with Dirichlet top, nfgs_solver.py:535-608."""
    Nz = rhs.shape[0]
    dz2 = dz * dz
    a = c = 1.0 / dz2
    b = (lam - 2.0 / dz2).astype(rhs.dtype)
    u = np.zeros_like(rhs)
    u[0] = u0
    u[Nz - 1] = utop
    n = Nz - 2
    if n <= 0:
        return u
    cs = np.empty((n,) + b.shape, dtype=rhs.dtype)
    ds = np.empty_like(cs)
    d0 = rhs[1] - a * u0
    if n == 1:
        cs[0] = 0.0
        ds[0] = (d0 - c * utop) / b
    else:
        cs[0] = c / b
        ds[0] = d0 / b
        for i in range(1, n):
            d = rhs[i + 1]
            ce = c
            if i == n - 1:
                d = d - c * utop
                ce = 0.0
            den = b - a * cs[i - 1]
            cs[i] = ce / den
            ds[i] = (d - a * ds[i - 1]) / den
    for i in range(n - 2, -1, -1):
        ds[i] = ds[i] - cs[i] * ds[i + 1]
    u[1:1 + n] = ds
    return u


def modal_robin_from_hat(src_hat, bottom_hat, lam, k, dz, rho=0.0, dtarget_hat=None, top_hat=None):
    """Modal (d_zz + lam) u = s, bottom Dirichlet (optionally rho-mixed first
    row) and eliminated top Robin row; nfgs_solver.py:5512-5625 (returns the
    modal solution, before the inverse transform)."""
    s = src_hat.copy()
    bh = bottom_hat.copy()
    s[:, 0, 0] = 0.0
    bh[0, 0] = 0.0
    g = np.zeros_like(bh) if top_hat is None else top_hat.copy()
    g[0, 0] = 0.0
    if rho > 0.0:
        q = dtarget_hat.copy()
        q[0, 0] = 0.0
    Nz = s.shape[0]
    dz2 = dz * dz
    lo = up = 1.0 / dz2
    diag = lam - 2.0 / dz2
    nu = Nz - 1
    cp = np.empty((nu,) + lam.shape, dtype=s.dtype)
    dp = np.empty_like(cp)
    if rho > 0.0:
        kb = (1.0 + 0.5 * rho) / dz2
        fd = (1.0 - rho) * lam - 2.0 / dz2
        fu = (1.0 - 0.5 * rho) / dz2
        fr = (1.0 - rho) * s[1]
        fr = fr - rho * q / dz
        r0 = fr - kb * bh
    else:
        fd, fu = diag, up
        r0 = s[1] - lo * bh
    cp[0] = fu / fd
    dp[0] = r0 / fd
    for r in range(1, nu - 1):
        den = diag - lo * cp[r - 1]
        cp[r] = up / den
        dp[r] = (s[r + 1] - lo * dp[r - 1]) / den
    rl = -lam * dz2 - 2.0
    rd = 2.0 + 2.0 * dz * k
    rr = 2.0 * dz * g - s[-2] * dz2
    last = nu - 1
    den = rd - rl * cp[last - 1]
    cp[last] = 0.0
    dp[last] = (rr - rl * dp[last - 1]) / den
    sol = np.empty_like(s)
    sol[0] = bh
    sol[-1] = dp[last]
    for r in range(last - 1, -1, -1):
        sol[r + 1] = dp[r] - cp[r] * sol[r + 2]
    sol[:, 0, 0] = 0.0
    return sol


# ----------------------------------------------------------------------------
# The solver restatement
# ----------------------------------------------------------------------------


@dataclass
class OracleResult:
    history: dict
    B: np.ndarray
    J: np.ndarray
    Phi_c: np.ndarray
    Az: np.ndarray
    extras: dict = field(default_factory=dict)


class OracleNFGS:
    """NumPy restatement of NFGSSolverGPU's default path (nfgs_solver.py:3905-7196)."""

    def __init__(self, Nx, Ny, Nz, dx=1.0, dy=1.0, dz=1.0, *, dtype=np.float64,
                 bottom_divergence_robin_weight=0.0, pad_calculation_domain=True,
                 calculation_domain_mode="mudpack", weak_field_floor_b2=1e-20,
                 alpha_limit_factor=1.0, source_zero_mode_policy="error",
                 source_zero_mode_tolerance=1e-2, d0_zero_mode_policy="error",
                 mean_mode_tolerance=1e-6, poloidal_update_mode="direct",
                 horizontal_zero_mode_policy="frozen_linear_harmonic", k0_ode_smoothness=1.0e-3,
                 coupled_k0_discretization="staggered_face", k0_update_mu=None,
                 k0_spline_basis_size=8, soft_cfcf_epsilon=None, soft_cfcf_normal_ridge=0.0):
        if poloidal_update_mode not in ("direct", "two_stage"):
            raise ValueError("poloidal_update_mode must be 'direct' or 'two_stage'")
        if horizontal_zero_mode_policy not in ("frozen_linear_harmonic", "coupled_k0",
                                               "constrained_spline_harmonic"):
            raise ValueError("unsupported horizontal_zero_mode_policy")
        self.poloidal_update_mode = poloidal_update_mode
        # coupled horizontal-harmonic closures (nfgs_solver.py:3963-3972,
        # :4120-4170, :4283-4287)
        self.k0_policy = horizontal_zero_mode_policy
        self.k0_smooth = float(k0_ode_smoothness)
        self.k0_disc = "legacy_plane_centred" if coupled_k0_discretization == \
            "legacy_plane_centred" else "staggered_face"
        if k0_update_mu is None and horizontal_zero_mode_policy == "constrained_spline_harmonic":
            k0_update_mu = 0.2
        self.k0_mu = None if k0_update_mu is None else float(k0_update_mu)
        self.k0_nb = int(k0_spline_basis_size)
        self.input_shape_zyx = (int(Nz), int(Ny), int(Nx))
        cx, cy, cz, pad_xy = calculation_grid(int(Nx), int(Ny), int(Nz), pad_calculation_domain,
                                              calculation_domain_mode)
        if cz < 3:
            raise ValueError("NFGS requires Nz >= 3")
        self.Nx, self.Ny, self.Nz = cx, cy, cz
        self.pad_xy = pad_xy
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.dtype = np.dtype(dtype)
        self.rho = float(bottom_divergence_robin_weight)
        self.floor_b2 = float(weak_field_floor_b2)
        self.alpha_max = float(alpha_limit_factor) / min(self.dx, self.dy, self.dz)
        self.src_policy = source_zero_mode_policy
        self.src_tol = float(source_zero_mode_tolerance)
        self.d0_policy = d0_zero_mode_policy
        self.mean_tol = float(mean_mode_tolerance)
        self.zero_mode_history = []
        # soft_cfcf start (nfgs_solver.py:3991-3996, :4218-4223, :4298-4305)
        h = min(self.dx, self.dy, self.dz)
        self.soft_eps = 1.0e-4 / (h * h) if soft_cfcf_epsilon is None else float(soft_cfcf_epsilon)
        self.soft_ridge = float(soft_cfcf_normal_ridge)
        if self.soft_eps < 0.0 or self.soft_ridge < 0.0:
            raise ValueError("soft_cfcf epsilon and ridge must be non-negative")

    # -- tables (nfgs_solver.py:1083-1118, :4357-4383) --
    def lam_neumann(self):
        kx = np.arange(self.Nx, dtype=self.dtype)
        ky = np.arange(self.Ny, dtype=self.dtype)
        lx = -4.0 * np.sin(np.pi * kx / (2.0 * max(self.Nx - 1, 1))) ** 2 / self.dx ** 2
        ly = -4.0 * np.sin(np.pi * ky / (2.0 * max(self.Ny - 1, 1))) ** 2 / self.dy ** 2
        return (ly[:, None] + lx[None, :]).astype(self.dtype)

    def lam_dirichlet(self):
        nxi, nyi = self.Nx - 2, self.Ny - 2
        kx = np.arange(1, nxi + 1, dtype=self.dtype)
        ky = np.arange(1, nyi + 1, dtype=self.dtype)
        lx = -4.0 * np.sin(np.pi * kx / (2.0 * (nxi + 1))) ** 2 / self.dx ** 2
        ly = -4.0 * np.sin(np.pi * ky / (2.0 * (nyi + 1))) ** 2 / self.dy ** 2
        return (ly[:, None] + lx[None, :]).astype(self.dtype)

    def alpha_x(self):
        kx = np.arange(self.Nx, dtype=self.dtype)
        return (2.0 * np.sin(np.pi * kx / (2.0 * max(self.Nx - 1, 1))) / self.dx).astype(self.dtype)

    def alpha_y(self):
        ky = np.arange(self.Ny, dtype=self.dtype)
        return (2.0 * np.sin(np.pi * ky / (2.0 * max(self.Ny - 1, 1))) / self.dy).astype(self.dtype)

    def _const_mode(self, hat):
        """nfgs_solver.py:4462-4468."""
        return hat[..., 0, 0] / (4.0 * max(self.Nx - 1, 1) * max(self.Ny - 1, 1))

    # -- spectral horizontal operators (nfgs_solver.py:4385-4457) --
    def spectral_dx_2d(self, v):
        hat = dct2(v)
        sine = -self.alpha_x() * hat
        tmp = idct1(sine, -2)
        out = np.zeros_like(v)
        if v.shape[-1] >= 3:
            out[..., 1:-1] = idst1(tmp[..., 1:-1], -1)
        return out

    def spectral_dy_2d(self, v):
        hat = dct2(v)
        sine = -self.alpha_y()[:, None] * hat
        tmp = idct1(sine, -1)
        out = np.zeros_like(v)
        if v.shape[-2] >= 3:
            out[..., 1:-1, :] = idst1(tmp[..., 1:-1, :], -2)
        return out

    def dstx_dcty_hat(self, bx):
        out = np.zeros_like(bx)
        if bx.shape[-1] < 3:
            return out
        tmp = dct1(bx, -2)
        out[..., 1:-1] = dst1(tmp[..., 1:-1], -1)
        return out

    def dctx_dsty_hat(self, by):
        out = np.zeros_like(by)
        if by.shape[-2] < 3:
            return out
        tmp = dct1(by, -1)
        out[..., 1:-1, :] = dst1(tmp[..., 1:-1, :], -2)
        return out

    def spectral_div_hat(self, bx, by):
        return self.alpha_x() * self.dstx_dcty_hat(bx) + self.alpha_y()[:, None] * self.dctx_dsty_hat(by)

    # -- 2-D Poisson (nfgs_solver.py:1132-1179) --
    def poisson2d_dirichlet(self, src):
        out = np.zeros_like(src)
        if self.Nx < 3 or self.Ny < 3:
            return out
        h = dst2(src[..., 1:-1, 1:-1])
        h = h / self.lam_dirichlet()
        out[..., 1:-1, 1:-1] = idst1(idst1(h, -1), -2)
        return out

    def poisson2d_neumann(self, src):
        h = dct2(src)
        ls = self.lam_neumann().copy()
        ls[0, 0] = 1.0
        h = h / ls
        h[..., 0, 0] = 0.0
        return idct1(idct1(h, -1), -2)

    # ------------------------------------------------------------------
    # Boundary setup (nfgs_solver.py:1999-2110, :3024-3057, :4548-4919)
    # ------------------------------------------------------------------
    def set_bottom_vector_boundary(self, Bx, By, Bz, ijzbcor=3):
        dt = self.dtype
        bx, by, bz = (np.asarray(a, dtype=dt) for a in (Bx, By, Bz))
        if bz.ndim == 3:
            bx, by, bz = bx[0], by[0], bz[0]
        if bz.shape != (self.Ny, self.Nx):
            obs_ny, obs_nx = bz.shape
            bz, fov = pad_or_crop_center_2d(bz, self.Ny, self.Nx)
            bx, _ = pad_or_crop_center_2d(bx, self.Ny, self.Nx)
            by, _ = pad_or_crop_center_2d(by, self.Ny, self.Nx)
            self.fov = fov
            self.observed_zyx = (min(self.Nz, self.input_shape_zyx[0]), obs_ny, obs_nx)
        else:
            self.fov = (0, self.Ny, 0, self.Nx)
            self.observed_zyx = (self.Nz, self.Ny, self.Nx)
        # padding dispatch of nfgs_solver.py:4503-4516
        calc = (self.Nz, self.Ny, self.Nx)
        padded = any(c > o for c, o in zip(calc, self.input_shape_zyx)) or \
            any(c > o for c, o in zip(calc, self.observed_zyx)) or any(p > 0 for p in self.pad_xy)
        if padded:
            self.bottom_current_grid_mode = "legacy_helper"
            jz = derive_bottom_jz_legacy(bx, by, self.dx, self.dy, self.Nz, ijzbcor)
        else:
            self.bottom_current_grid_mode = "calculation_grid"
            jz = derive_bottom_jz_calc(bx, by, self.dx, self.dy, ijzbcor)
        self.Bx_obs, self.By_obs, self.Bz_obs, self.Jz_obs = bx, by, bz, jz.astype(dt)
        # Az_bottom and the toroidal bottom current (nfgs_solver.py:2105, :3037-3056)
        self.Az_bottom = self.poisson2d_dirichlet(-self.Jz_obs)
        jzt = -lap_xy(self.Az_bottom, self.dx, self.dy)
        jzt[:, 0] = 0.0
        jzt[:, -1] = 0.0
        jzt[0, :] = 0.0
        jzt[-1, :] = 0.0
        self.Jz_T = jzt
        rms = lambda v: float(np.sqrt(np.mean(np.asarray(v, dtype=np.float64) ** 2)))
        self.toroidal_diagnostics = {
            "raw_jz_rms": float(np.sqrt(np.mean(self.Jz_obs ** 2))),
            "t_implied_jz_rms": float(np.sqrt(np.mean(jzt ** 2))),
            "raw_vs_t_implied_jz_rms": float(np.sqrt(np.mean((self.Jz_obs - jzt) ** 2))),
        }
        self._gaussian_background()
        self.zero_mode_history = []
        self.last_zero_mode_by_height = []

    def _gaussian_background(self):
        """nfgs_solver.py:4671-4919 (uniform_open, spectral operator, no external field)."""
        dt = self.dtype
        bxh, byh, bzh = dct2(self.Bx_obs), dct2(self.By_obs), dct2(self.Bz_obs)
        self.B_mean = np.array([self._const_mode(bxh), self._const_mode(byh), self._const_mode(bzh)])
        rbz_hat = dct2(self.Bz_obs)
        self.Bz_open_mean = dt.type(self._const_mode(rbz_hat))
        flux_imb = float(abs(self.Bz_obs.sum()) / max(float(np.abs(self.Bz_obs).sum()), 1e-30))
        nonmean = rbz_hat.copy()
        nonmean[0, 0] = 0.0
        d_hat = self.spectral_div_hat(self.Bx_obs, self.By_obs)
        self.D0 = idct2(d_hat)
        d0_mean = self._const_mode(d_hat)
        d0_rms = np.sqrt(np.mean(self.D0 ** 2))
        d0_ratio = float(abs(d0_mean) / max(d0_rms, 1e-30))
        self._zero_mode_policy(d0_ratio, self.mean_tol, self.d0_policy,
                               "The bottom vector field is incompatible with the zero-mean "
                               "endpoint-Neumann GS model")
        d_hat[0, 0] = 0.0
        lam = self.lam_neumann()
        k = np.sqrt(np.clip(-lam, 0.0, None)).astype(dt)
        ks = k.copy()
        ks[0, 0] = 1.0
        ls = lam.copy()
        ls[0, 0] = 1.0
        self.k = k
        dphi_total_hat = -d_hat / ls
        dphi_total_hat[0, 0] = 0.0
        bzb = 0.5 * (nonmean + d_hat / ks)
        bza = 0.5 * (nonmean - d_hat / ks)
        bzb[0, 0] = 0.0
        bza[0, 0] = 0.0
        pbh = bzb / ls
        pah = bza / ls
        pbh[0, 0] = 0.0
        pah[0, 0] = 0.0
        self.phi_below_hat, self.phi_above_hat = pbh, pah
        self.Bz_below = idct2(bzb)
        self.Bz_above = idct2(bza)
        self.Phi_below_bottom = idct2(pbh)
        self.Phi_above_bottom = idct2(pah)
        z = np.arange(self.Nz, dtype=dt) * dt.type(self.dz)
        decay = np.exp(-z[:, None, None] * k[None, :, :])
        phi_below = idct2(decay * pbh[None])
        bz_below = idct2(decay * bzb[None])
        dphi_dz = idct2((-k)[None] * decay * pbh[None])
        gx, gy = self.spectral_dx_2d(dphi_dz), self.spectral_dy_2d(dphi_dz)
        bg = np.stack([-gx, -gy, bz_below])
        bg[2] = bg[2] + self.Bz_open_mean  # B_below + B_open (+ zero external)
        self.B_background = bg.astype(dt)
        self.dphi_target = (idct2(dphi_total_hat) - d_dz(phi_below[:3], self.dz)[0]).astype(dt)
        recon = self.B_background[2, 0] + self.Bz_above - self.Bz_obs
        b_from_phi = np.stack([-gx, -gy, lap_xy(phi_below, self.dx, self.dy)])
        b_from_phi[2] = b_from_phi[2] + self.Bz_open_mean
        inv = b_from_phi - self.B_background
        self.gaussian_diagnostics = {
            "bottom_reconstruction_rms": float(np.sqrt(np.mean(recon ** 2))),
            "bottom_reconstruction_max": float(np.abs(recon).max()),
            "dct_zero_mode_Bx": float(self.B_mean[0]),
            "dct_zero_mode_By": float(self.B_mean[1]),
            "dct_zero_mode_Bz": float(self.B_mean[2]),
            "Bz_open_mean": float(self.Bz_open_mean),
            "flux_imbalance_fraction": flux_imb,
            "D0_zero_mode": float(d0_mean),
            "D0_zero_mode_ratio": d0_ratio,
            "d0_removed_constant_mode": float(d0_mean),
            "background_scalar_plus_explicit_rms": float(np.sqrt(np.mean(inv ** 2))),
        }
        self.gaussian_diagnostics.update(self._harmonic_diagnostics())
        # every closure starts from the bottom target repeated with height
        # (nfgs_solver.py:5058-5096, :5150-5190)
        self.H = np.repeat(self.H_target[:, None], self.Nz, axis=1).astype(dt)
        self.k0_residual_history = []
        self.k0_diag = {}
        if self.k0_policy != "frozen_linear_harmonic":
            self._k0_setup()

    def _harmonic_diagnostics(self):
        """nfgs_solver.py:4990-5056."""
        dphi = idct2(self.k * (self.phi_above_hat - self.phi_below_hat))
        gx, gy = self.spectral_dx_2d(dphi), self.spectral_dy_2d(dphi)
        btx = d_dy(self.Az_bottom, self.dy)
        bty = -d_dx(self.Az_bottom, self.dx)
        rec = np.stack([-gx + btx, -gy + bty])
        obs = np.stack([self.Bx_obs, self.By_obs])
        res = obs - rec
        const = np.array([self._const_mode(dct2(res[0])), self._const_mode(dct2(res[1]))],
                         dtype=self.dtype)
        self.H_target = const
        nonconst = res - const[:, None, None]
        div = idct2(self.spectral_div_hat(res[0], res[1]))
        curl = d_dx(res[1], self.dx) - d_dy(res[0], self.dy)
        rms = lambda v: float(np.sqrt(np.mean(v ** 2)))
        ratio = rms(res) / max(rms(obs), 1e-30)
        return {
            "bottom_horizontal_reconstruction_rms": rms(res),
            "bottom_horizontal_reconstruction_ratio": ratio,
            "harmonic_residual_constant_rms": rms(const),
            "harmonic_residual_nonconstant_rms": rms(nonconst),
            "harmonic_residual_divergence_rms": rms(div),
            "harmonic_residual_curl_rms": rms(curl),
        }

    def gaussian_above_bottom_field(self):
        """The outward Gaussian field on z = 0, (3, Ny, Nx) (nfgs_solver.py:4921-4943):
        B_h = -spectral grad_h of idct2(k Phi^>_hat), B_z = Bz_above."""
        dphi = idct2(self.k * self.phi_above_hat)
        gx, gy = self.spectral_dx_2d(dphi), self.spectral_dy_2d(dphi)
        return np.stack([-gx, -gy, self.Bz_above]).astype(self.dtype)

    def _zero_mode_policy(self, ratio, tol, policy, message):
        """nfgs_solver.py:4477-4487."""
        if ratio <= tol:
            return
        detail = f"{message} ratio={ratio:.6e} exceeds tolerance={tol:.6e}."
        if policy == "error":
            raise ValueError(detail)
        if policy == "warn_project":
            import warnings
            warnings.warn(detail + " The incompatible zero mode will be projected out.",
                          RuntimeWarning, stacklevel=3)

    # ------------------------------------------------------------------
    # Scalar solves (nfgs_solver.py:3263-3306, :5457-5476, :5627-5672)
    # ------------------------------------------------------------------
    def diagnose_zero_mode(self, src):
        means = self._const_mode(dct2(src))
        src_rms = np.sqrt(np.mean(src ** 2))
        zero_rms = np.sqrt(np.mean(means ** 2))
        ratio = float(zero_rms / max(src_rms, 1e-30))
        self.last_zero_mode_by_height = [float(v) for v in means]
        self.zero_mode_history.append(ratio)
        self._zero_mode_policy(ratio, self.src_tol, self.src_policy,
                               "The Phi_c nonlinear source has an incompatible zero mode")
        return ratio

    def solve_phi_c(self, src):
        """The poloidal update dispatch (nfgs_solver.py:5727-5733)."""
        if self.poloidal_update_mode == "two_stage":
            return self.solve_phi_c_two_stage(src)
        return self.solve_phi_c_direct(src)

    def solve_phi_c_two_stage(self, src):
        """B_c,z from the modal Robin solve with the source undivided and the
        bottom data scaled by lambda, then Phi_c_hat = DCT2(B_c,z) / lambda
        plane by plane (nfgs_solver.py:5674-5725)."""
        lam = self.lam_neumann()
        ls = lam.copy()
        ls[0, 0] = 1.0
        self.diagnose_zero_mode(src)
        sh = dct2(src)
        sh[:, 0, 0] = 0.0
        bh = dct2(self.Phi_above_bottom)
        bh[0, 0] = 0.0
        qh = None
        if self.rho > 0.0:
            qh = dct2(self.dphi_target)
            qh[0, 0] = 0.0
            qh = lam * qh
        bcz = idct2(modal_robin_from_hat(sh, lam * bh, lam, self.k, self.dz, self.rho, qh))
        ph = dct2(bcz) / ls[None]
        ph[:, 0, 0] = 0.0
        phi = idct2(ph)
        phi[0] = self.Phi_above_bottom
        return phi

    def solve_phi_c_direct(self, src):
        lam = self.lam_neumann()
        ls = lam.copy()
        ls[0, 0] = 1.0
        self.diagnose_zero_mode(src)
        sh = dct2(src) / ls[None]
        sh[:, 0, 0] = 0.0
        bh = dct2(self.Phi_above_bottom)
        bh[0, 0] = 0.0
        qh = None
        if self.rho > 0.0:
            qh = dct2(self.dphi_target)
            qh[0, 0] = 0.0
        sol = modal_robin_from_hat(sh, bh, lam, self.k, self.dz, self.rho, qh)
        phi = idct2(sol)
        if self.rho > 0.0:
            phi[0] = self.Phi_above_bottom
        return phi

    def solve_az(self, src):
        Az = np.zeros_like(src)
        Az[0] = self.Az_bottom
        Az[-1] = 0.0
        lam = self.lam_dirichlet()
        sh = dst2(src[:, 1:-1, 1:-1])
        rh = -sh / lam
        bh = dst2(self.Az_bottom[1:-1, 1:-1])
        th = np.zeros_like(bh)
        ah = thomas_dirichlet(rh, bh, th, lam, self.dz)
        Az[:, 1:-1, 1:-1] = idst2(ah)
        Az[0] = self.Az_bottom
        Az[-1] = 0.0
        self._az_lateral_zero(Az)
        return Az

    @staticmethod
    def _az_lateral_zero(Az):
        Az[:, :, 0] = 0.0
        Az[:, :, -1] = 0.0
        Az[:, 0, :] = 0.0
        Az[:, -1, :] = 0.0

    # ------------------------------------------------------------------
    # Staggered-face current (nfgs_solver.py:5738-5839)
    # ------------------------------------------------------------------
    def face_state(self, B, Jz):
        dz = self.dz
        Bf = 0.5 * (B[:, 1:] + B[:, :-1])
        c = B - self.B_background
        cf = 0.5 * (c[:, 1:] + c[:, :-1])
        dbx = (c[0, 1:] - c[0, :-1]) / dz
        dby = (c[1, 1:] - c[1, :-1]) / dz
        jx = d_dy(cf[2], self.dy) - dby
        jy = dbx - d_dx(cf[2], self.dx)
        jz = 0.5 * (Jz[1:] + Jz[:-1])
        return Bf, np.stack([jx, jy, jz])

    @staticmethod
    def faces_to_centres(v):
        out = np.empty((v.shape[0], v.shape[1] + 1) + v.shape[2:], dtype=v.dtype)
        out[:, 0] = v[:, 0]
        out[:, -1] = v[:, -1]
        out[:, 1:-1] = 0.5 * (v[:, :-1] + v[:, 1:])
        return out

    def parallel_current(self, J, B):
        Bf, Jf = self.face_state(B, J[2])
        b2 = np.maximum(np.sum(Bf * Bf, axis=0), self.floor_b2)
        alpha = np.sum(Jf * Bf, axis=0) / b2
        alpha = np.clip(alpha, -self.alpha_max, self.alpha_max)
        return self.faces_to_centres(alpha[None] * Bf)

    # ------------------------------------------------------------------
    # Metrics (nfgs_solver.py:2656-2697, :5882-5895)
    # ------------------------------------------------------------------
    def window(self, nz_faces):
        nzo, nyo, nxo = self.observed_zyx
        bfy, bfx = self.fov[0], self.fov[2]
        bfz = max(0, nz_faces - int(nzo))
        ax, ay, az = nxo // 8 + bfx, nyo // 8 + bfy, nzo // 4 + bfz
        sz = slice(1, max(1, nz_faces - 1 - az + 1))
        sy = slice(ay, max(ay + 1, self.Ny - 1 - ay + 1))
        sx = slice(ax, max(ax + 1, self.Nx - 1 - ax + 1))
        return sz, sy, sx

    def sigma(self, B, J):
        Bf, Jf = self.face_state(B, J[2])
        sz, sy, sx = self.window(Bf.shape[1])
        b = Bf[:, sz, sy, sx]
        j = Jf[:, sz, sy, sx]
        bn = np.maximum(np.sqrt(np.sum(b * b, axis=0)), 1e-10)
        jn = np.sqrt(np.sum(j * j, axis=0))
        jxb = np.cross(np.moveaxis(j, 0, -1), np.moveaxis(b, 0, -1))
        num = (np.sqrt(np.sum(jxb * jxb, axis=-1)) / bn).sum()
        den = max(jn.sum(), 1e-30)
        return float(num / den)

    def divB(self, B):
        div = d_dx(B[0], self.dx) + d_dy(B[1], self.dy) + d_dz(B[2], self.dz)
        bm = np.sqrt(B[0] ** 2 + B[1] ** 2 + B[2] ** 2)
        delta = min(self.dx, self.dy, self.dz)
        return float(np.abs(div).mean() / (6.0 * bm.mean() / delta + 1e-12))

    def energy(self, B):
        return float((B[0] ** 2 + B[1] ** 2 + B[2] ** 2).sum() * self.dx * self.dy * self.dz / 2.0)

    # ------------------------------------------------------------------
    # Reconstruction, init and one step (nfgs_solver.py:3311-3341,
    # :5901-5942, :6069-6078, :6613-6935)
    # ------------------------------------------------------------------
    # ------------------------------------------------------------------
    # Coupled horizontal-harmonic closures (coupled_k0 and
    # constrained_spline_harmonic): the uniform field H(z) follows the
    # horizontal mean of the parallel current not carried by the PT scalars
    # (nfgs_solver.py:5058-5398, :5952-6067, :6760-6791), in FP64.
    # ------------------------------------------------------------------
    @staticmethod
    def bspline_basis(n_points, n_basis):
        """Clamped open-uniform cubic B-spline collocation (nfgs_solver.py:5098-5147)."""
        deg = 3
        inner = n_basis - deg - 1
        knot = [0.0] * (deg + 1) + [i / (inner + 1) for i in range(1, inner + 1)] + [1.0] * (deg + 1)
        knot = np.array(knot)
        t = np.linspace(0.0, 1.0, n_points)
        width = knot.size - 1
        basis = np.zeros((n_points, width))
        for i in range(width):
            basis[:, i] = ((t >= knot[i]) & (t < knot[i + 1])).astype(float)
        basis[-1, -1] = 1.0
        for order in range(1, deg + 1):
            nxt = np.zeros((n_points, width - 1))
            for i in range(width - 1):
                ld = knot[i + order] - knot[i]
                rd = knot[i + order + 1] - knot[i + 1]
                if ld > 0.0:
                    nxt[:, i] += (t - knot[i]) / ld * basis[:, i]
                if rd > 0.0:
                    nxt[:, i] += (knot[i + order + 1] - t) / rd * basis[:, i + 1]
            basis = nxt
            width -= 1
            basis[-1] = 0.0
            basis[-1, -1] = 1.0
        return basis

    def _k0_setup(self):
        nz, dz = self.Nz, self.dz
        s = self.k0_smooth
        D = np.zeros((nz - 1, nz))
        r = np.arange(nz - 1)
        D[r, r] = -1.0 / dz
        D[r, r + 1] = 1.0 / dz
        C = np.zeros((max(nz - 2, 0), nz))
        if nz >= 3:
            r = np.arange(nz - 2)
            C[r, r] = 1.0 / dz
            C[r, r + 1] = -2.0 / dz
            C[r, r + 2] = 1.0 / dz
        if self.k0_policy == "constrained_spline_harmonic":
            basis = self.bspline_basis(nz, self.k0_nb)
            pieces = [D @ basis]
            if s > 0.0 and C.shape[0] > 0:
                pieces.append(np.sqrt(s) * (C @ basis))
            design = np.concatenate(pieces, axis=0)
            normal = design.T @ design
            ridge = max(np.abs(np.diagonal(normal)).max(), 1.0) * 1.0e-12
            normal = normal + ridge * np.eye(self.k0_nb)
            top = np.zeros(nz)
            top[-1], top[-2], top[-3] = 3.0 / (2 * dz), -4.0 / (2 * dz), 1.0 / (2 * dz)
            cons = np.stack([basis[0], top @ basis])
            kkt = np.block([[normal, cons.T], [cons, np.zeros((2, 2))]])
            self.k0_mats = dict(basis=basis, design=design, kkt=kkt, C=C)
            self.k0_diag = {"enabled": True, "closure": "constrained_spline_harmonic",
                            "basis_size": self.k0_nb, "smoothness": s,
                            "update_mu": self.k0_mu,
                            "kkt_condition_number": float(np.linalg.cond(kkt))}
            return
        if self.k0_disc == "legacy_plane_centred":
            Dl = np.zeros((nz, nz))
            r = np.arange(1, nz - 1)
            Dl[r, r - 1] = -0.5 / dz
            Dl[r, r + 1] = 0.5 / dz
            Dl[0, :3] = [-1.5 / dz, 2.0 / dz, -0.5 / dz]
            Dl[nz - 1, nz - 3:] = [0.5 / dz, -2.0 / dz, 1.5 / dz]
            Cl = Dl @ Dl
            pieces = [Dl[1:, 1:]]
            if s > 0.0:
                pieces.append(np.sqrt(s) * Cl[1:, 1:])
            D, C = Dl, Cl
        else:
            pieces = [D[:, 1:]]
            if C.shape[0] > 0 and s > 0.0:
                pieces.append(np.sqrt(s) * C[:, 1:])
        A = np.concatenate(pieces, axis=0)
        self.k0_mats = dict(A=A, L=np.linalg.cholesky(A.T @ A), D=D, C=C)
        self.k0_diag = {"enabled": True, "smoothness": s, "discretization": self.k0_disc}

    def _k0_profile(self, q, h0):
        """One component of the closure's profile (nfgs_solver.py:5229-5261, :5355-5398)."""
        m = self.k0_mats
        s = np.sqrt(self.k0_smooth)
        if self.k0_policy == "constrained_spline_harmonic":
            rhs = [0.5 * (q[:-1] + q[1:])]
            if self.k0_smooth > 0.0 and m["C"].shape[0] > 0:
                rhs.append(np.zeros(m["C"].shape[0]))
            nrhs = m["design"].T @ np.concatenate(rhs)
            sol = np.linalg.solve(m["kkt"], np.concatenate([nrhs, [h0, 0.0]]))
            return m["basis"] @ sol[:self.k0_nb]
        D, C = m["D"], m["C"]
        if self.k0_disc == "legacy_plane_centred":
            rhs = [q[1:] - D[1:, 0] * h0]
            if self.k0_smooth > 0.0:
                rhs.append(-s * C[1:, 0] * h0)
        else:
            rhs = [0.5 * (q[:-1] + q[1:]) - D[:, 0] * h0]
            if C.shape[0] > 0 and self.k0_smooth > 0.0:
                rhs.append(-s * C[:, 0] * h0)
        nrhs = m["A"].T @ np.concatenate(rhs)
        L = m["L"]
        y = np.linalg.solve(L, nrhs)
        return np.concatenate([[h0], np.linalg.solve(L.T, y)])

    def _k0_centre_derivative(self, p):
        """nfgs_solver.py:5400-5431 (either discretisation)."""
        if self.k0_policy == "coupled_k0" and self.k0_disc == "legacy_plane_centred":
            return (self.k0_mats["D"] @ p.T).T
        d = np.zeros_like(p)
        face = (p[:, 1:] - p[:, :-1]) / self.dz
        d[:, 1:-1] = 0.5 * (face[:, :-1] + face[:, 1:])
        d[:, 0] = (-3.0 * p[:, 0] + 4.0 * p[:, 1] - p[:, 2]) / (2.0 * self.dz)
        d[:, -1] = (3.0 * p[:, -1] - 4.0 * p[:, -2] + p[:, -3]) / (2.0 * self.dz)
        return d

    def k0_centre_derivative_of(self, p):
        return self._k0_centre_derivative(np.asarray(p, dtype=np.float64)).astype(self.dtype)

    def _trap_mean(self, f):
        """The constant DCT-I mode of every plane (nfgs_solver.py:4466-4468)."""
        return dct2(f)[:, 0, 0] / (4.0 * (self.Nx - 1) * (self.Ny - 1))

    def k0_update(self, jp, phi, az, step_mu=1.0):
        """The closure's profile from the step's parallel current and the new
        scalars (nfgs_solver.py:5952-6067, :6760-6791); sets self.H.  The
        relaxation is k0_update_mu, or the step's mu without it (:6776)."""
        bcz = lap_xy(phi, self.dx, self.dy)
        _, jpt, _ = reconstruct_pt(phi, az, bcz, self.B_background, self.Bz_obs, self.Jz_T,
                                   self.dx, self.dy, self.dz)
        want = np.stack([self._trap_mean(jp[0]) - self._trap_mean(jpt[0]),
                         self._trap_mean(jp[1]) - self._trap_mean(jpt[1])]).astype(np.float64)
        hb = self.H_target.astype(np.float64)
        prof = np.stack([self._k0_profile(want[1], hb[0]), self._k0_profile(-want[0], hb[1])])
        d = self._k0_centre_derivative(prof)
        actual = np.stack([-d[1], d[0]])
        res = actual - want
        rel = np.sqrt(np.mean(res ** 2)) / max(np.sqrt(np.mean(want ** 2)), 1e-30)
        self.k0_residual_history.append(rel)
        face = (prof[:, 1:] - prof[:, :-1]) / self.dz
        want_face = np.stack([0.5 * (want[1, :-1] + want[1, 1:]),
                              -0.5 * (want[0, :-1] + want[0, 1:])])
        fres = face - want_face
        grid = prof[:, 1:-1] - 0.25 * (prof[:, :-2] + 2.0 * prof[:, 1:-1] + prof[:, 2:])
        prms = np.sqrt(np.mean(prof ** 2))
        self.k0_diag.update(
            ode_residual_rms=float(np.sqrt(np.mean(res ** 2))), ode_residual_relative=float(rel),
            ode_residual_relative_max=float(max(self.k0_residual_history)),
            face_ode_residual_rms=float(np.sqrt(np.mean(fres ** 2))),
            face_ode_residual_relative=float(np.sqrt(np.mean(fres ** 2)) /
                                             max(np.sqrt(np.mean(want_face ** 2)), 1e-30)),
            gridscale_profile_rms=float(np.sqrt(np.mean(grid ** 2))),
            gridscale_profile_relative=float(np.sqrt(np.mean(grid ** 2)) / max(prms, 1e-30)),
            top_derivative=[float(v) for v in d[:, -1]],
            profile_max_abs=float(np.abs(prof).max()),
            profile_top=[float(v) for v in prof[:, -1]],
            desired_mean_current_rms=float(np.sqrt(np.mean(want ** 2))))
        mu = float(step_mu) if self.k0_mu is None else self.k0_mu
        new = prof if mu >= 1.0 else mu * prof + (1.0 - mu) * self.H.astype(np.float64)
        self.H = new.astype(self.dtype)
        self.H_unrelaxed = prof
        return prof

    def _apply_bc(self, B, J):
        J[:, :, :, 0] = 0.0
        J[:, :, :, -1] = 0.0
        J[:, :, 0, :] = 0.0
        J[:, :, -1, :] = 0.0
        J[:, -1] = 0.0
        B[2, 0] = self.Bz_obs
        J[2, 0] = self.Jz_T

    def k0_current(self):
        """nfgs_solver.py:5400-5455 (centre derivative of the constant profile)."""
        p = self.H
        d = np.zeros_like(p)
        face = (p[:, 1:] - p[:, :-1]) / self.dz
        d[:, 1:-1] = 0.5 * (face[:, :-1] + face[:, 1:])
        d[:, 0] = (-3.0 * p[:, 0] + 4.0 * p[:, 1] - p[:, 2]) / (2.0 * self.dz)
        d[:, -1] = (3.0 * p[:, -1] - 4.0 * p[:, -2] + p[:, -3]) / (2.0 * self.dz)
        return d

    def reconstruct(self, Phi_c, Az, Bc_z):
        dx, dy, dz = self.dx, self.dy, self.dz
        dphi = d_dz(Phi_c, dz)
        bcx = -d_dx(dphi, dx) + d_dy(Az, dy)
        bcy = -d_dy(dphi, dy) - d_dx(Az, dx)
        jz = -lap_xy(Az, dx, dy)
        B = self.B_background + np.stack([bcx, bcy, Bc_z])
        # Jx/Jy of the PT reconstruction (nfgs_solver.py:1760-1768) are
        # replaced by the staggered-face current below, so zeros stand in.
        J = np.stack([np.zeros_like(jz), np.zeros_like(jz), jz])
        self._apply_bc(B, J)
        d = self.k0_current()
        B[0] = B[0] + self.H[0][:, None, None]
        B[1] = B[1] + self.H[1][:, None, None]
        B[2, 0] = self.Bz_obs
        J[2, 0] = self.Jz_T
        J[:, -1] = 0.0
        del d  # contributes only to Jx/Jy, which are overwritten
        Bf, Jf = self.face_state(B, J[2])
        Js = self.faces_to_centres(Jf)
        if self.k0_policy == "constrained_spline_harmonic":
            # walls close the PT current only; the uniform harmonic current
            # stays (nfgs_solver.py:5919-5932)
            dk = self.k0_centre_derivative_of(self.H)
            jh = np.zeros_like(Js)
            jh[0] = -dk[1][:, None, None]
            jh[1] = dk[0][:, None, None]
            jpt = Js - jh
            jpt[:, :, :, 0] = 0.0
            jpt[:, :, :, -1] = 0.0
            jpt[:, :, 0, :] = 0.0
            jpt[:, :, -1, :] = 0.0
            jpt[:, -1] = 0.0
            J = jpt + jh
            J[:, -1] = 0.0
        else:
            J[:2] = Js[:2]
            J[:, :, :, 0] = 0.0
            J[:, :, :, -1] = 0.0
            J[:, :, 0, :] = 0.0
            J[:, :, -1, :] = 0.0
            J[:, -1] = 0.0
        J[2, 0] = self.Jz_T
        self.last_face_diag = {
            "face_current_rms": float(np.sqrt(np.mean(Jf.astype(np.float64) ** 2))),
            "projected_cell_current_rms": float(np.sqrt(np.mean(Js.astype(np.float64) ** 2))),
        }
        return B, J

    def initial_state(self):
        z = np.zeros((self.Nz, self.Ny, self.Nx), dtype=self.dtype)
        phi = self.solve_phi_c(z)
        az = self.solve_az(z)
        B, J = self.reconstruct(phi, az, lap_xy(phi, self.dx, self.dy))
        return B, J, phi, az

    # ------------------------------------------------------------------
    # soft_cfcf start (nfgs_solver.py:6152-6611)
    # ------------------------------------------------------------------
    def _soft_profile_map(self, k, kind):
        """The z profile of every mode as p = E u + e b: u the free values,
        b the bottom value.  The bottom value is exact; for Phi_c the second
        plane makes the bottom derivative equal k b (:6244-6256), and the top
        plane closes the transparent Robin row (3 p_N - 4 p_{N-1} + p_{N-2}) /
        (2 dz) + k p_N = 0 (:6266-6271).  k: (M,).  Returns E (M, N, n) and
        e (M, N)."""
        N, dz, M = self.Nz, self.dz, k.shape[0]
        first = 2 if kind == "phi" else 1  # plane of the first free value
        n = N - 1 - first
        E = np.zeros((M, N, n))
        e = np.zeros((M, N))
        e[:, 0] = 1.0
        for z in range(first, N - 1):
            E[:, z, z - first] = 1.0
        if kind == "phi":
            E[:, 1, 0] = 0.25
            e[:, 1] = 0.75 + 0.5 * dz * k
        den = 3.0 + 2.0 * dz * k
        E[:, N - 1] = (4.0 / den)[:, None] * E[:, N - 2] - (1.0 / den)[:, None] * E[:, N - 3]
        e[:, N - 1] = (4.0 / den) * e[:, N - 2] - (1.0 / den) * e[:, N - 3]
        return E, e

    def _soft_operators(self):
        """Second derivative on the interior planes and the first derivative
        of an interior sequence (one-sided second order at its ends), as
        matrices in z (:6178-6238)."""
        N, dz = self.Nz, self.dz
        D2 = np.zeros((N - 2, N))
        for r in range(N - 2):
            D2[r, r:r + 3] = np.array([1.0, -2.0, 1.0]) / (dz * dz)
        m = N - 2
        D1 = np.zeros((m, m))
        D1[0, :3] = np.array([-3.0, 4.0, -1.0]) / (2.0 * dz)
        for r in range(1, m - 1):
            D1[r, r - 1], D1[r, r + 1] = -1.0 / (2.0 * dz), 1.0 / (2.0 * dz)
        D1[-1, -3:] = np.array([1.0, -4.0, 3.0]) / (2.0 * dz)
        Dc = np.zeros((N - 2, N))  # centred first derivative at the interior planes
        for r in range(N - 2):
            Dc[r, r], Dc[r, r + 2] = -1.0 / (2.0 * dz), 1.0 / (2.0 * dz)
        return D2, D1, Dc

    def soft_modes(self, k, bottom, kind):
        """Modal profiles minimising the modal form of int |curl J_c|^2 +
        eps int |J_c|^2 (:6322-6402) with the essential rows exact: for
        Phi_c |d_z L_k p|^2 + (k^2 + eps) |L_k p|^2, for Az |L_k p|^2 +
        eps (|d_z p|^2 + k^2 |p|^2), L_k = d_zz - k^2 on the interior
        planes.  The normal equations of each mode are solved densely (the
        reference factors the band); a ridge adds ridge * max(1, max|H_ii|)
        to the diagonal.  k, bottom: arrays of modes; modes with k <= 0 give
        zero.  Returns (N,) + k.shape profiles in FP64."""
        k = np.asarray(k, dtype=np.float64)
        b = np.asarray(bottom, dtype=np.float64)
        shape = k.shape
        kf, bf = k.reshape(-1), b.reshape(-1)
        out = np.zeros((self.Nz, kf.size))
        act = np.nonzero(kf > 0.0)[0]
        D2, D1, Dc = self._soft_operators()
        N = self.Nz
        eps = self.soft_eps
        for start in range(0, act.size, 2048):
            idx = act[start:start + 2048]
            kk, bb = kf[idx], bf[idx]
            E, e = self._soft_profile_map(kk, kind)
            k2 = kk * kk
            # L_k on the interior planes, per mode: (M, N-2, N)
            Lk = D2[None] - k2[:, None, None] * np.eye(N)[None, 1:N - 1]
            if kind == "phi":
                blocks = [(np.einsum("ij,mjk->mik", D1, Lk), np.ones_like(kk)),
                          (Lk, k2 + eps)]
            else:
                blocks = [(Lk, np.ones_like(kk))]
                if eps > 0.0:
                    blocks += [(np.broadcast_to(Dc, (kk.size,) + Dc.shape), np.full_like(kk, eps)),
                               (np.broadcast_to(np.eye(N)[1:N - 1], (kk.size, N - 2, N)),
                                eps * k2)]
            n = E.shape[2]
            H = np.zeros((kk.size, n, n))
            g = np.zeros((kk.size, n))
            for R, w in blocks:
                A = np.einsum("mij,mjk->mik", R, E)          # rows on the free values
                c = np.einsum("mij,mj->mi", R, e) * bb[:, None]  # rows' constants
                H += w[:, None, None] * np.einsum("mri,mrj->mij", A, A)
                g += w[:, None] * np.einsum("mri,mr->mi", A, c)
            if self.soft_ridge > 0.0:
                dmax = np.maximum(np.abs(np.diagonal(H, axis1=1, axis2=2)).max(axis=1), 1.0)
                H[:, np.arange(n), np.arange(n)] += (self.soft_ridge * dmax)[:, None]
            u = np.linalg.solve(H, -g[..., None])[..., 0]
            out[:, idx] = (np.einsum("mzi,mi->mz", E, u) + e * bb[:, None]).T
        return out.reshape((self.Nz,) + shape)

    def soft_initial_state(self):
        """Phi_c and Az of the soft_cfcf start and their state
        (:6423-6484): every DCT mode of Phi_c and every DST mode of Az
        from soft_modes, the bottom planes and the lateral Az walls fixed
        afterwards.  Returns B, J, Phi_c, Az and the modal profiles."""
        if self.Nz < 5:
            raise ValueError("soft_cfcf requires Nz >= 5")
        dt = self.dtype
        pbh = dct2(self.Phi_above_bottom)
        pbh[0, 0] = 0.0
        phi_hat = self.soft_modes(self.k, pbh, "phi").astype(dt)
        phi = idct2(phi_hat).astype(dt)
        phi[0] = self.Phi_above_bottom
        abh = dst1(dst1(self.Az_bottom[1:-1, 1:-1], -1), -2)
        ak = np.sqrt(np.clip(-self.lam_dirichlet(), 0.0, None)).astype(dt)
        az_hat = self.soft_modes(ak, abh, "az").astype(dt)
        az = np.zeros((self.Nz, self.Ny, self.Nx), dtype=dt)
        az[:, 1:-1, 1:-1] = idst1(idst1(az_hat, -1), -2)
        az[0] = self.Az_bottom
        B, J = self.reconstruct(phi, az, lap_xy(phi, self.dx, self.dy))
        self.soft_profiles = (phi_hat, az_hat, ak)
        return B, J, phi, az

    def soft_diagnostics(self, B, J, phi, az):
        """The soft_cfcf diagnostics (:6486-6611): the modal energies of the
        profiles, the reconstructed state's curl and current energies, the
        bottom rows and the transparent top rows."""
        phi_hat, az_hat, ak = (np.asarray(a, dtype=np.float64) for a in self.soft_profiles)
        dz = self.dz
        kp = np.asarray(self.k, dtype=np.float64)
        q = (phi_hat[2:] - 2.0 * phi_hat[1:-1] + phi_hat[:-2]) / (dz * dz) - kp ** 2 * phi_hat[1:-1]
        dq = np.empty_like(q)
        dq[0] = (-3.0 * q[0] + 4.0 * q[1] - q[2]) / (2.0 * dz)
        dq[1:-1] = (q[2:] - q[:-2]) / (2.0 * dz)
        dq[-1] = (3.0 * q[-1] - 4.0 * q[-2] + q[-3]) / (2.0 * dz)
        w = np.ones((self.Ny, self.Nx))
        w[[0, -1], :] *= 0.5
        w[:, [0, -1]] *= 0.5
        phi_cur = kp ** 2 * (q ** 2).sum(axis=0)
        phi_curl = kp ** 2 * (dq ** 2 + kp ** 2 * q ** 2).sum(axis=0)
        qa = (az_hat[2:] - 2.0 * az_hat[1:-1] + az_hat[:-2]) / (dz * dz) - ak ** 2 * az_hat[1:-1]
        da = (az_hat[2:] - az_hat[:-2]) / (2.0 * dz)
        az_cur = ak ** 2 * (da ** 2 + ak ** 2 * az_hat[1:-1] ** 2).sum(axis=0)
        az_curl = ak ** 2 * (qa ** 2).sum(axis=0)
        f = self.dx * self.dy * dz / (4.0 * (self.Nx - 1) * (self.Ny - 1))
        curl_e = f * ((w * phi_curl).sum() + az_curl.sum())
        cur_e = f * ((w * phi_cur).sum() + az_cur.sum())
        Jd = {"Bx": J[0], "By": J[1], "Bz": J[2]}
        cj = fd2_curl(Jd, (self.dx, self.dy, dz))
        vol = self.dx * self.dy * dz
        rec_curl = float(sum((v.astype(np.float64) ** 2).sum() for v in cj.values()) * vol)
        rec_cur = float((J.astype(np.float64) ** 2).sum() * vol)
        rms = lambda v: float(np.sqrt(np.mean(np.asarray(v, dtype=np.float64) ** 2)))
        target = idct2(self.k * dct2(self.Phi_above_bottom))
        dphi0 = fd2_derivative(phi, dz, 0)[0]
        ph = dct2(phi)
        top_phi = idct2((3.0 * ph[-1] - 4.0 * ph[-2] + ph[-3]) / (2.0 * dz) + self.k * ph[-1])
        ah = dst1(dst1(az[:, 1:-1, 1:-1], -1), -2)
        ak_t = np.sqrt(np.clip(-self.lam_dirichlet(), 0.0, None)).astype(self.dtype)
        top_az = idst1(idst1((3.0 * ah[-1] - 4.0 * ah[-2] + ah[-3]) / (2.0 * dz)
                             + ak_t * ah[-1], -1), -2)
        bc = np.stack([-d_dx(target, self.dx) + d_dy(self.Az_bottom, self.dy),
                       -d_dy(target, self.dy) - d_dx(self.Az_bottom, self.dx)])
        dp0 = d_dz(phi[:3], dz)[0]
        bcx0 = -d_dx(dp0, self.dx) + d_dy(az[0], self.dy)
        bcy0 = -d_dy(dp0, self.dy) - d_dx(az[0], self.dx)
        return {
            "curl_current_energy": float(curl_e), "current_energy": float(cur_e),
            "objective": float(curl_e + self.soft_eps * cur_e),
            "reconstructed_curl_current_energy": rec_curl,
            "reconstructed_current_energy": rec_cur,
            "bottom_phi_rms": rms(phi[0] - self.Phi_above_bottom),
            "bottom_dphi_rms": rms(dphi0 - target),
            "bottom_az_rms": rms(az[0] - self.Az_bottom),
            "bottom_bz_rms": rms(B[2, 0] - self.Bz_obs),
            "bottom_jz_rms": rms(J[2, 0] - self.Jz_T),
            "bottom_bh_current_carrying_rms": rms(np.stack([bcx0, bcy0]) - bc),
            "bottom_bh_observation_rms": rms(B[:2, 0] - np.stack([self.Bx_obs, self.By_obs])),
            "top_phi_transparent_rms": rms(top_phi),
            "top_az_transparent_rms": rms(top_az),
        }

    def project_field(self, B, Az_in):
        """Caller-field start: the cfcf handoff of nfgs_solver.py:6080-6147 with
        (B, Az) given.  Phi_c = skip-bottom endpoint-Neumann Poisson of
        Bc_z = B_z - B_bg,z (:1120-1130, :1146-1180), the constant mode of every
        plane by height (:4466-4468), Az with the solver's fixed planes; the
        projected state's B and J by reconstruct().  Returns
        (Phi_c, Az, mean_bcz_by_height, B_projected, J_projected)."""
        bcz = B[2] - self.B_background[2]
        mean = dct2(bcz)[:, 0, 0] / (4.0 * (self.Nx - 1) * (self.Ny - 1))
        phi = np.empty_like(bcz)
        phi[0] = self.Phi_above_bottom
        phi[1:] = self.poisson2d_neumann(bcz[1:])
        az = np.array(Az_in, dtype=self.dtype, copy=True)
        az[0] = self.Az_bottom
        az[-1] = 0.0
        self._az_lateral_zero(az[1:])
        Bp, Jp = self.reconstruct(phi, az, lap_xy(phi, self.dx, self.dy))
        return phi, az, mean, Bp, Jp

    def sources(self, B, J, polarity="both"):
        """Parallel current, bottom row, and both Poisson sources
        (nfgs_solver.py:6694-6745).  polarity "positive" ("negative") keeps
        the bottom row only where the observed B_z is positive (negative)
        (:6696-6701)."""
        jp = self.parallel_current(J, B)
        jp[:, 0] = J[:, 0]
        if polarity != "both":
            keep = (self.Bz_obs > 0) if polarity == "positive" else (self.Bz_obs < 0)
            jp[:, 0] = J[:, 0] * keep.astype(self.dtype)
        s_bcz = d_dy(jp[0], self.dy) - d_dx(jp[1], self.dx)
        s_jz = lap_xy(jp[2], self.dx, self.dy) - d2_dxdz(jp[0], self.dx, self.dz) \
            - d2_dydz(jp[1], self.dy, self.dz)
        return jp, s_bcz, s_jz

    def step(self, B, J, Phi_c, Az, mu=1.0, polarity="both"):
        """One map of the state (nfgs_solver.py:6647-6805).  With mu < 1 the
        new scalars are mu times the solves' plus (1 - mu) times the passed
        ones, before the fixed planes are set again (:6778-6790)."""
        jp, s_bcz, s_jz = self.sources(B, J, polarity)
        phi = self.solve_phi_c(s_bcz)
        az = self.solve_az(s_jz)
        phi[0] = self.Phi_above_bottom
        az[0] = self.Az_bottom
        az[-1] = 0.0
        self._az_lateral_zero(az)
        if self.k0_policy != "frozen_linear_harmonic":
            self.k0_update(jp, phi, az, mu)
        if mu < 1.0:
            phi = mu * phi + (1.0 - mu) * Phi_c
            az = mu * az + (1.0 - mu) * Az
            phi[0] = self.Phi_above_bottom
            az[0] = self.Az_bottom
            az[-1] = 0.0
            self._az_lateral_zero(az)
        B2, J2 = self.reconstruct(phi, az, lap_xy(phi, self.dx, self.dy))
        return B2, J2, phi, az

    # ------------------------------------------------------------------
    # Per-iteration diagnostics (record_iteration_diagnostics=True,
    # nfgs_solver.py:6660-6680 and :6806-6931), at mu = 1 with the frozen
    # horizontal zero mode and no bottom projection
    # ------------------------------------------------------------------
    DIAGNOSTIC_KEYS = ("alpha_clip_fraction", "az_map_residual_relative",
                       "az_update_relative", "bcz_pde_residual_relative",
                       "bottom_field_imprint_relative", "bottom_jz_imprint_relative",
                       "h_k0_map_residual_relative", "h_k0_update_relative",
                       "jz_pde_residual_relative", "mixed_row_residual_relative",
                       "phi_map_residual_relative", "phi_update_relative",
                       "poloidal_pde_residual_relative", "toroidal_pde_residual_relative",
                       "weak_field_floor_fraction")

    def step_diagnostics(self, B, J, Phi_c, Az):
        """One step and its diagnostics: returns (B2, J2, phi, az, diag)."""
        dx, dy, dz = self.dx, self.dy, self.dz
        f64 = lambda v: np.asarray(v, dtype=np.float64)
        rms = lambda v: float(np.sqrt(np.mean(f64(v) ** 2)))

        def rel(delta, scale):
            return rms(delta) / max(rms(scale), 1.0e-30)

        def rel_op(residual, *terms):
            return rms(residual) / max(max(rms(t) for t in terms), 1.0e-30)

        # face counts of the input state (:6660-6680)
        Bf, Jf = self.face_state(B, J[2])
        b2 = np.sum(Bf * Bf, axis=0)
        weak = float(np.mean(f64(b2 <= self.floor_b2)))
        alpha = np.sum(Jf * Bf, axis=0) / np.maximum(b2, self.dtype.type(self.floor_b2))
        clip = float(np.mean(f64(np.abs(alpha) > self.alpha_max)))

        _, s_bcz, s_jz = self.sources(B, J)
        B2, J2, phi, az = self.step(B, J, Phi_c, Az)
        lap = lambda f: lap_xy(f, dx, dy)
        d2z = lambda f: d2_dz2(f, dz)
        bcz = lap(phi)
        bcz_res = d2z(bcz) + lap(bcz) - s_bcz
        jzs = -lap(az)
        jz_res = d2z(jzs) + lap(jzs) - s_jz
        zb = slice(2, -1) if self.Nz > 4 else slice(1, -1)
        zj = slice(1, -1)

        lam = self.lam_neumann().astype(self.dtype)
        ls = lam.copy()
        ls[0, 0] = 1.0
        sph = dct2(s_bcz) / ls[None]
        sph[:, 0, 0] = 0.0
        src_phi = idct2(sph)
        phi_res = d2z(phi) + lap(phi) - src_phi

        src_az = np.zeros_like(az)
        sjh = dst2(s_jz[:, 1:-1, 1:-1])
        src_az[:, 1:-1, 1:-1] = idst2(-sjh / self.lam_dirichlet().astype(self.dtype)[None])
        az_res = d2z(az) + lap(az) - src_az
        inner = (slice(1, -1), slice(1, -1), slice(1, -1))

        rho = self.rho
        dphi = (-3.0 * phi[0] + 4.0 * phi[1] - phi[2]) / (2.0 * dz)
        mixed_pde = (1.0 - rho) * phi_res[1]
        mixed_der = -rho * (dphi - self.dphi_target) / dz
        mixed = mixed_pde + mixed_der

        obs = np.stack([self.Bx_obs, self.By_obs, self.Bz_obs])
        diag = {
            "phi_update_relative": rel(phi - Phi_c, Phi_c),
            "az_update_relative": rel(az - Az, Az),
            "phi_map_residual_relative": rel(phi - Phi_c, Phi_c),
            "az_map_residual_relative": rel(az - Az, Az),
            "h_k0_update_relative": 0.0,  # the frozen profile does not change in a step
            "h_k0_map_residual_relative": 0.0,
            "bcz_pde_residual_relative": rel(bcz_res[zb], s_bcz[zb]),
            "jz_pde_residual_relative": rel(jz_res[zj], s_jz[zj]),
            "poloidal_pde_residual_relative": rel_op(phi_res[zb], d2z(phi)[zb], lap(phi)[zb],
                                                     src_phi[zb]),
            "toroidal_pde_residual_relative": rel_op(az_res[inner], d2z(az)[inner],
                                                     lap(az)[inner], src_az[inner]),
            "mixed_row_residual_relative": rel_op(mixed, mixed_pde, mixed_der, src_phi[1]),
            "bottom_field_imprint_relative": rel(B2[:, 0] - obs, obs),
            "bottom_jz_imprint_relative": rel(J2[2, 0] - self.Jz_T, self.Jz_T),
            "alpha_clip_fraction": clip,
            "weak_field_floor_fraction": weak,
        }
        return B2, J2, phi, az, diag

    # ------------------------------------------------------------------
    # Driver (nfgs_solver.py:3522-3850, :6937-7009)
    # ------------------------------------------------------------------
    def iterate(self, max_iter=200, tol=1e-5, patience=1, sigma_improvement_tol=1e-5,
                callback=None, start=None, record_diagnostics=False, stop_mode="sigma_best",
                fixed_point_update_tol=None, fixed_point_pde_tol=1.0e-8,
                fixed_point_observable_tol=None, fixed_point_patience=10, mu=1.0,
                adaptive_mu=False, mu_min=0.02, mu_max=1.0, mu_backoff=0.5, mu_growth=1.1,
                growth_streak=5, divergence_factor=2.0, max_rollbacks=12, polarity="both"):
        """start=(B, Az): begin from project_field(B, Az) instead of gs_cfcf.
        record_diagnostics: the per-step diagnostics as history lists;
        stop_mode="fixed_point" (needs them) stops after fixed_point_patience
        consecutive steps with the map residual, the PDE residual and the
        sigma/energy drift within their tolerances and selects the final
        iterate (nfgs_solver.py:3565-3582, :3712-3733, :3766-3776, :3816).
        mu < 1 relaxes every step; adaptive_mu rolls a step whose sigma is
        not finite or above divergence_factor times the best back to the best
        state with mu backed off, grows mu after growth_streak new bests, and
        stops after max_rollbacks (:3658-3762); polarity masks the bottom
        parallel current."""
        if record_diagnostics and (mu != 1.0 or adaptive_mu or polarity != "both"):
            raise ValueError("per-step diagnostics are restated at mu = 1 only")
        if stop_mode == "fixed_point" and not record_diagnostics:
            raise ValueError("stop_mode='fixed_point' requires record_iteration_diagnostics=True")
        upd = float(tol if fixed_point_update_tol is None else fixed_point_update_tol)
        obs_tol = float(tol if fixed_point_observable_tol is None else fixed_point_observable_tol)
        pde_tol = float(fixed_point_pde_tol)
        streak = 0
        fp_converged = False
        diags = []
        if start is None:
            B, J, phi, az = self.initial_state()
        elif isinstance(start, str) and start == "soft_cfcf":
            B, J, phi, az = self.soft_initial_state()
        else:
            phi, az, _, B, J = self.project_field(*start)
        s0, d0, e0 = self.sigma(B, J), self.divB(B), self.energy(B)
        h = {k: [] for k in ("sigma", "divB", "energy", "mu", "rollback", "sigma_relative_change",
                             "energy_relative_change", "fixed_point_criteria_met")}
        best_sigma, best_iter = s0, 0
        best = (B.copy(), J.copy(), phi.copy(), az.copy())
        best_h = self.H.copy()
        stall = flat = 0
        need = max(int(patience), 1)
        prev = s0
        use_imp = sigma_improvement_tol is not None and float(sigma_improvement_tol) > 0.0
        eps = float(sigma_improvement_tol) if use_imp else 0.0
        last_imp = None
        reason = "max_iter"
        final_iter = 0
        mu_cur = float(mu)
        rollbacks = improve = 0
        for n in range(1, max_iter + 1):
            if record_diagnostics:
                Bn, Jn, pn, an, dg = self.step_diagnostics(B, J, phi, az)
                diags.append(dg)
            else:
                Bn, Jn, pn, an = self.step(B, J, phi, az, mu_cur, polarity)
            s = self.sigma(Bn, Jn)
            if adaptive_mu and (not np.isfinite(s) or s > divergence_factor * max(best_sigma, 1e-30)):
                rollbacks += 1
                mu_cur = max(mu_cur * mu_backoff, mu_min)
                B, J, phi, az = (a.copy() for a in best)
                self.H = best_h.copy()
                improve = 0
                for key, v in (("sigma", s), ("divB", np.nan), ("energy", np.nan), ("mu", mu_cur),
                               ("rollback", 1), ("sigma_relative_change", np.nan),
                               ("energy_relative_change", np.nan),
                               ("fixed_point_criteria_met", 0)):
                    h[key].append(v)
                streak = 0
                if rollbacks >= max_rollbacks:
                    reason = "max_rollbacks"
                    break
                continue
            B, J, phi, az = Bn, Jn, pn, an
            d, e = self.divB(B), self.energy(B)
            pe = h["energy"][-1] if h["energy"] else e0
            src = abs(s - prev) / max(abs(prev), 1e-30)
            erc = abs(e - pe) / max(abs(pe), 1e-30)
            h["sigma"].append(s)
            h["divB"].append(d)
            h["energy"].append(e)
            h["mu"].append(mu_cur)
            h["rollback"].append(0)
            h["sigma_relative_change"].append(src)
            h["energy_relative_change"].append(erc)
            met = False
            if stop_mode == "fixed_point":
                dg = diags[-1]
                mres = max(dg["phi_map_residual_relative"], dg["az_map_residual_relative"],
                           dg["h_k0_map_residual_relative"])
                pres = max(dg["poloidal_pde_residual_relative"],
                           dg["toroidal_pde_residual_relative"], dg["mixed_row_residual_relative"])
                met = (np.isfinite(mres) and np.isfinite(pres) and mres <= upd and pres <= pde_tol
                       and src <= obs_tol and erc <= obs_tol)
            h["fixed_point_criteria_met"].append(int(met))
            streak = streak + 1 if met else 0
            final_iter = n
            if callback is not None:
                callback(n, s, B, J)
            imp = prev - s
            last_imp = imp
            prev = s
            flat = flat + 1 if imp < eps else 0
            if s < best_sigma:
                best_sigma, best_iter = s, n
                best = (B.copy(), J.copy(), phi.copy(), az.copy())
                best_h = self.H.copy()
                stall = 0
                improve += 1
                if adaptive_mu and improve >= growth_streak:
                    mu_cur = min(mu_cur * mu_growth, mu_max)
                    improve = 0
            else:
                stall += 1
                improve = 0
            if stop_mode == "fixed_point":
                if streak >= int(fixed_point_patience):
                    reason = "fixed_point"
                    fp_converged = True
                    break
                continue
            if s < tol:
                reason = "tol"
                break
            if use_imp and flat >= need:
                reason = "sigma_improvement_tol"
                break
            if patience > 0 and stall >= patience:
                reason = "sigma_stall"
                break
        final = (B, J)
        final_state = (B, J, phi, az)
        fvb = np.sqrt(np.mean((final[0].astype(np.float64) - best[0].astype(np.float64)) ** 2)) / \
            max(np.sqrt(np.mean(best[0].astype(np.float64) ** 2)), 1e-30)
        if record_diagnostics:
            for k in self.DIAGNOSTIC_KEYS:
                h[k] = [float(dg[k]) for dg in diags]
        if stop_mode == "fixed_point":
            best = tuple(a.copy() for a in final_state)
            best_h = self.H.copy()
        # the selected state's harmonic profile (nfgs_solver.py:6990-7004)
        self.H = best_h
        h.update(fixed_point_converged=fp_converged, fixed_point_streak=streak)
        h.update(best_sigma=best_sigma, best_iter=best_iter, rollbacks=rollbacks, stop_reason=reason,
                 stop_mode=stop_mode,
                 selected_state="final_accepted_iterate" if stop_mode == "fixed_point"
                 else "best_sigma", final_iter=final_iter,
                 final_sigma=float(self.sigma(*final)), final_vs_best_B_relative=float(fvb),
                 sigma_improvement_tol=eps if use_imp else None, last_sigma_improvement=last_imp,
                 initial_sigma=s0, initial_divB=d0, initial_energy=e0)
        self.B, self.J, self.Phi_c, self.Az = best
        # post-selection staggered diagnostic (nfgs_solver.py:7005-7007)
        Bf, Jf = self.face_state(self.B, self.J[2])
        Js = self.faces_to_centres(Jf)
        self.last_face_diag = {
            "face_current_rms": float(np.sqrt(np.mean(Jf.astype(np.float64) ** 2))),
            "projected_cell_current_rms": float(np.sqrt(np.mean(Js.astype(np.float64) ** 2))),
        }
        h["source_diagnostics"] = self.source_diagnostics(self.B)
        return h

    def fields_to_numpy(self, crop_to_observed_fov=True):
        """nfgs_solver.py:2881-2903."""
        B, J = self.B, self.J
        if crop_to_observed_fov:
            nzo = self.observed_zyx[0]
            y0, y1, x0, x1 = self.fov
            B = B[:, :nzo, y0:y1, x0:x1]
            J = J[:, :nzo, y0:y1, x0:x1]
        return {"Bx": B[0].copy(), "By": B[1].copy(), "Bz": B[2].copy(),
                "Jx": J[0].copy(), "Jy": J[1].copy(), "Jz": J[2].copy()}

    # ------------------------------------------------------------------
    # Source diagnostics (nfgs_solver.py:7014-7196; numeric entries)
    # ------------------------------------------------------------------
    def source_diagnostics(self, B):
        dz = self.dz
        bcz = B[2] - self.B_background[2]
        recon = self.B_background[2, 0] + self.Bz_above - self.Bz_obs
        corr = bcz[0] - self.Bz_above
        bc_hat = dct2(bcz[-3:])
        der = (3.0 * bc_hat[-1] - 4.0 * bc_hat[-2] + bc_hat[-3]) / (2.0 * dz)
        rh = der + self.k * bc_hat[-1]
        rh[0, 0] = 0.0
        robin = idct2(rh)
        phi = np.empty_like(bcz)
        phi[0] = self.Phi_above_bottom
        phi[1:] = self.poisson2d_neumann(bcz[1:])
        ph = dct2(phi[-3:])
        pder = (3.0 * ph[-1] - 4.0 * ph[-2] + ph[-3]) / (2.0 * dz)
        prh = pder + self.k * ph[-1]
        prh[0, 0] = 0.0
        probin = idct2(prh)
        ks = self.k.copy()
        ks[0, 0] = 1.0
        uh = 0.5 * (bc_hat[-1] + der / ks)
        uh[0, 0] = 0.0
        up = idct2(uh)
        rms = lambda v: float(np.sqrt(np.mean(np.asarray(v) ** 2)))
        div = d_dx(B[0, 0], self.dx) + d_dy(B[1, 0], self.dy)
        derr = div - self.D0
        dscale = max(np.sqrt(np.mean(self.D0 ** 2)), 1e-30)
        dphi = (-3.0 * phi[0] + 4.0 * phi[1] - phi[2]) / (2.0 * dz)
        tor = self.toroidal_diagnostics
        return {
            "bottom_reconstruction_rms": rms(recon),
            "bottom_reconstruction_max": float(np.abs(recon).max()),
            "correction_bottom_rms": rms(corr),
            "top_robin_rms": rms(robin),
            "top_phi_robin_rms": rms(probin),
            "top_upward_source_rms": rms(up),
            "top_upward_source_ratio": float(rms(up) / max(rms(bcz[-1]), 1e-30)),
            "source_zero_mode_ratio_last": self.zero_mode_history[-1] if self.zero_mode_history else 0.0,
            "source_zero_mode_ratio_max": max(self.zero_mode_history) if self.zero_mode_history else 0.0,
            "bottom_divergence_rms": rms(derr),
            "bottom_divergence_relative": float(np.sqrt(np.mean(derr ** 2)) / dscale),
            "bottom_dphi_residual_rms": rms(dphi - self.dphi_target),
            "Jz_projected_out_rms": tor["raw_vs_t_implied_jz_rms"],
            "B_harm_over_Bobs_rms": self.gaussian_diagnostics["bottom_horizontal_reconstruction_ratio"],
        }


# ----------------------------------------------------------------------
# Common scoring of output fields (FP64): the second-order curl and
# divergence of nfgs/diagnostics/common_fd2.py:23-61, sigma_J on a box
# (common_fd2.py:64-83, util/metrics.py:66-92) and the reference-field
# agreement metrics (util/metrics.py:55-63, :128-176).
# ----------------------------------------------------------------------
def fd2_derivative(f, h, axis):
    """Centred inside, second-order one-sided on the end samples; an axis of 2
    samples takes the forward difference at both, an axis of 1 gives 0."""
    g = np.moveaxis(np.asarray(f, dtype=np.float64), axis, 0)
    d = np.zeros_like(g)
    n = g.shape[0]
    if n == 2:
        d[0] = d[1] = (g[1] - g[0]) / h
    elif n >= 3:
        d[1:-1] = (g[2:] - g[:-2]) / (2.0 * h)
        d[0] = (4.0 * g[1] - 3.0 * g[0] - g[2]) / (2.0 * h)
        d[-1] = (3.0 * g[-1] - 4.0 * g[-2] + g[-3]) / (2.0 * h)
    return np.moveaxis(d, 0, axis)


def fd2_curl(fields, spacing):
    """Curl of z-y-x arrays; spacing is (dx, dy, dz) or one number."""
    dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
    bx, by, bz = fields["Bx"], fields["By"], fields["Bz"]
    return {"Jx": fd2_derivative(bz, dy, 1) - fd2_derivative(by, dz, 0),
            "Jy": fd2_derivative(bx, dz, 0) - fd2_derivative(bz, dx, 2),
            "Jz": fd2_derivative(by, dx, 2) - fd2_derivative(bx, dy, 1)}


def fd2_divergence(fields, spacing):
    dx, dy, dz = (spacing,) * 3 if np.isscalar(spacing) else spacing
    return (fd2_derivative(fields["Bx"], dx, 2) + fd2_derivative(fields["By"], dy, 1)
            + fd2_derivative(fields["Bz"], dz, 0))


def box_cwsin(B, J, b2min=1e-10):
    """sum |J x B| / max(|B|, b2min) over sum |J| on component-first boxes."""
    B = np.asarray(B, dtype=np.float64)
    J = np.asarray(J, dtype=np.float64)
    c = np.stack([J[1] * B[2] - J[2] * B[1], J[2] * B[0] - J[0] * B[2],
                  J[0] * B[1] - J[1] * B[0]])
    bn = np.sqrt(np.sum(B * B, axis=0))
    num = np.sum(np.sqrt(np.sum(c * c, axis=0)) / np.maximum(bn, b2min))
    return float(num / max(float(np.sum(np.sqrt(np.sum(J * J, axis=0)))), 1e-30))


def common_sigma_j(fields, spacing, slices):
    j = fd2_curl(fields, spacing)
    sz, sy, sx = slices
    B = np.stack([fields[k] for k in ("Bx", "By", "Bz")])[:, sz, sy, sx]
    J = np.stack([j[k] for k in ("Jx", "Jy", "Jz")])[:, sz, sy, sx]
    return box_cwsin(B, J)


def reference_metrics(ref_fields, fields, slices):
    sz, sy, sx = slices
    R = np.stack([ref_fields[k] for k in ("Bx", "By", "Bz")]).astype(np.float64)[:, sz, sy, sx]
    S = np.stack([fields[k] for k in ("Bx", "By", "Bz")]).astype(np.float64)[:, sz, sy, sx]
    J = np.stack([fields[k] for k in ("Jx", "Jy", "Jz")]).astype(np.float64)[:, sz, sy, sx]
    rn = np.sqrt(np.sum(R * R, axis=0))
    sn = np.sqrt(np.sum(S * S, axis=0))
    dot = np.sum(R * S, axis=0)
    en = np.sqrt(np.sum((S - R) ** 2, axis=0))
    c_vec = float(np.sum(dot)) / (float(np.sqrt(np.sum(S * S) * np.sum(R * R))) + 1e-12)
    e_n = float(np.sum(en) / (np.sum(rn) + 1e-12))
    e_m = float(np.mean(en / (rn + 1e-12)))
    return {"C_vec": c_vec, "C_CS": float(np.mean(dot / (rn * sn + 1e-12))), "En": e_n, "Em": e_m,
            "Eprime_n": 1.0 - e_n, "Eprime_m": 1.0 - e_m,
            "epsilon": float(np.sum(S * S) / (np.sum(R * R) + 1e-12)), "sigma_J": box_cwsin(S, J)}


def td_metrics(reference, candidate, current, divergence, slices, spacing):
    """The controlled evaluation's metrics
    (nfgs/_backend/scripts/recompute_td_referee_diagnostics.py:344-400): the
    agreement metrics and sigma_J of reference_metrics plus the divergence
    normalised by 6 mean|B| / spacing."""
    sz, sy, sx = slices
    fields = dict(candidate)
    fields.update({k: current[k] for k in ("Jx", "Jy", "Jz")})
    m = reference_metrics(reference, fields, slices)
    S = np.stack([candidate[k] for k in ("Bx", "By", "Bz")]).astype(np.float64)[:, sz, sy, sx]
    bn_mean = float(np.mean(np.sqrt(np.sum(S * S, axis=0))))
    div_mean = float(np.mean(np.abs(np.asarray(divergence, dtype=np.float64)[sz, sy, sx])))
    return {k: m[k] for k in ("C_vec", "C_CS", "Eprime_n", "Eprime_m", "epsilon", "sigma_J")} | {
        "divB_normalized": div_mean / (6.0 * bn_mean / float(spacing) + 1e-12)}


def outward_errors(target, candidate, slices, cutoffs):
    """Errors of an outward bottom field against a target on the y/x window of
    `slices` (nfgs/controlled.py:55-68): ||c - b|| / ||b|| over the window
    (None for a zero target) and, per cutoff fraction, the mean of
    |c - b| / |b| over the pixels with |b| > cutoff * max|b| (None when no
    pixel is kept), with the kept count and fraction."""
    _, sy, sx = slices
    b = np.asarray(target, dtype=np.float64)[:, sy, sx]
    c = np.asarray(candidate, dtype=np.float64)[:, sy, sx]
    mag = np.sqrt(np.sum(b * b, axis=0))
    err = np.sqrt(np.sum((c - b) ** 2, axis=0))
    tot = float(np.sqrt(np.sum(b * b)))
    res = {"E_rel_Bgt_L2": float(np.sqrt(np.sum((c - b) ** 2)) / tot) if tot > 0 else None,
           "CCI_pixels_total": int(mag.size), "CCI_cutoff_results": {}}
    for cut in cutoffs:
        keep = mag > float(cut) * float(mag.max())
        res["CCI_cutoff_results"][str(cut)] = {
            "E_rel_Bgt": float(np.mean(err[keep] / mag[keep])) if keep.any() else None,
            "retained_pixels": int(keep.sum()), "retained_fraction": float(keep.mean())}
    res["E_rel_Bgt"] = res["CCI_cutoff_results"][str(cutoffs[0])]["E_rel_Bgt"]
    return res
