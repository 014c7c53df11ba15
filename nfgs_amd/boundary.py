"""Seeded synthetic photospheric boundaries for the benchmark workloads.

The reference ships no generator for these inputs, so this module defines
them: Gaussian flux poles for Bz and a constant-alpha linear force-free
field for the horizontal components, computed with an FFT on the observed
grid in FP64.  With k the horizontal wavevector and l = sqrt(k^2 - alpha^2)
the decaying solution's bottom values are

    Bx_hat = i (alpha k_y - l k_x) / k^2 * Bz_hat
    By_hat = -i (alpha k_x + l k_y) / k^2 * Bz_hat

(the Nakagawa-Raadu / Alissandrakis form); modes with k = 0 or k^2 <= alpha^2
have no decaying solution and are set to zero.
"""
from __future__ import annotations

import numpy as np


def lfff_horizontal(bz: np.ndarray, alpha: float, dx: float = 1.0, dy: float = 1.0):
    """Horizontal field of the constant-alpha linear force-free extrapolation."""
    ny, nx = bz.shape
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dy)
    KY, KX = np.meshgrid(ky, kx, indexing="ij")
    k2 = KX * KX + KY * KY
    ok = k2 > alpha * alpha
    l = np.sqrt(np.where(ok, k2 - alpha * alpha, 0.0))
    safe = np.where(ok, k2, 1.0)
    bzh = np.fft.fft2(bz)
    bxh = np.where(ok, 1j * (alpha * KY - l * KX) / safe, 0.0) * bzh
    byh = np.where(ok, -1j * (alpha * KX + l * KY) / safe, 0.0) * bzh
    return np.real(np.fft.ifft2(bxh)), np.real(np.fft.ifft2(byh))


def _gauss(nyo, nxo, x0, y0, sig):
    y, x = np.meshgrid(np.arange(nyo, dtype=np.float64), np.arange(nxo, dtype=np.float64),
                       indexing="ij")
    return np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2.0 * sig * sig))


def dipole_boundary(n: int = 64, seed: int = 0, alpha: float = 0.02, sigma: float = 6.0,
                    separation: float = 20.0, noise: float = 1e-3):
    """Two Gaussian poles (+1/-1), centred pair at angle U[0, pi), noise on Bz."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0.0, np.pi)
    c = 0.5 * (n - 1)
    ox, oy = 0.5 * separation * np.cos(th), 0.5 * separation * np.sin(th)
    bz = _gauss(n, n, c + ox, c + oy, sigma) - _gauss(n, n, c - ox, c - oy, sigma)
    bx, by = lfff_horizontal(bz, alpha)
    bz_obs = bz + rng.normal(0.0, noise, size=bz.shape)
    return bx, by, bz_obs


def multipole_boundary(nx: int = 256, ny: int | None = None, seed: int = 0,
                       npoles: tuple = (4, 8), amp: tuple = (0.5, 2.0), sigma: tuple = (8.0, 20.0),
                       alpha_max: float = 0.03, drift: float = 0.0, alpha_scale: float = 1.0):
    """4-8 flux-balanced Gaussian poles and a random-alpha LFFF horizontal field.

    Returns (Bx, By, Bz, alpha).  Negative poles are rescaled so that the net
    flux sum(a_i sigma_i^2) vanishes.  drift > 0 moves every pole by that many
    cells in a seeded direction and alpha_scale scales alpha: a later frame of
    the same region (drift=0, alpha_scale=1 give the base frame).
    """
    ny = nx if ny is None else ny
    rng = np.random.default_rng(seed)
    n = int(rng.integers(npoles[0], npoles[1] + 1))
    a = rng.uniform(amp[0], amp[1], n) * rng.choice([-1.0, 1.0], n)
    if np.all(a > 0):
        a[0] = -a[0]
    if np.all(a < 0):
        a[0] = -a[0]
    s = rng.uniform(sigma[0], sigma[1], n)
    cx = np.array([rng.uniform(3 * si, nx - 3 * si) for si in s])
    cy = np.array([rng.uniform(3 * si, ny - 3 * si) for si in s])
    flux = a * s * s
    pos, neg = flux[flux > 0].sum(), -flux[flux < 0].sum()
    a = np.where(a < 0, a * pos / neg, a)
    alpha = float(rng.uniform(-alpha_max, alpha_max)) * alpha_scale
    if drift:
        th = np.random.default_rng(seed + 7919).uniform(0.0, 2.0 * np.pi, n)
        cx = cx + drift * np.cos(th)
        cy = cy + drift * np.sin(th)
    bz = sum(ai * _gauss(ny, nx, xi, yi, si) for ai, xi, yi, si in zip(a, cx, cy, s))
    bx, by = lfff_horizontal(bz, alpha)
    return bx, by, bz, alpha
