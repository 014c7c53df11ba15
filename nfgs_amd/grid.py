"""Calculation-grid sizing, field-of-view embedding and the sigma window.

Integer host logic only.  Follows nfgs_solver.py:73-177 (MUDPACK sizes),
:771-819 (calculation-domain modes), :1593-1609 (centred embedding) and
:2656-2676 (the Fortran cwsin window).
"""
from __future__ import annotations

import math
from dataclasses import dataclass


def mudpack_axis_size(n_obs: int) -> int:
    """Smallest 2*2^p+1 or 3*2^p+1 size >= n_obs (nfgs_solver.py:73-83)."""
    n_obs = int(n_obs)
    if n_obs < 2:
        return max(1, n_obs)
    n = float(n_obs)
    e2 = math.ceil(math.log((n - 1.0) / 2.0) / math.log(2.0)) + 1
    e3 = math.ceil(math.log((n - 1.0) / 3.0) / math.log(2.0)) + 1
    return min(2 * 2 ** (e2 - 1) + 1, 3 * 2 ** (e3 - 1) + 1)


def _even_pad(n_calc: int, n_obs: int) -> int:
    """Grow by one so left/right pads match (nfgs_solver.py:86-102)."""
    if n_calc < n_obs:
        raise ValueError(f"calculation size {n_calc} is smaller than observed size {n_obs}")
    return n_calc + 1 if (n_calc - n_obs) % 2 == 1 else n_calc


@dataclass(frozen=True)
class HelperGrid:
    nx: int
    ny: int
    nz: int
    bfx: int
    bfy: int


def fortran_helper(nxo: int, nyo: int, nzo: int, symmetric_xy: bool = False) -> HelperGrid:
    """MUDPACK sizes and low-side pads of nfgs_solver.py:105-177."""
    nx, ny, nz = mudpack_axis_size(nxo), mudpack_axis_size(nyo), mudpack_axis_size(nzo)
    if symmetric_xy:
        nx, ny = _even_pad(nx, nxo), _even_pad(ny, nyo)
    return HelperGrid(nx, ny, nz, (nx - nxo) // 2, (ny - nyo) // 2)


def calculation_grid(Nx: int, Ny: int, Nz: int, pad: bool, mode: str):
    """(Nx, Ny, Nz, (pad_x, pad_y)) actually solved (nfgs_solver.py:771-819)."""
    mode = str(mode).lower().replace("-", "_")
    if mode not in ("mudpack", "mudpack_legacy", "bottom_padding"):
        raise ValueError(
            "calculation_domain_mode must be 'mudpack', 'mudpack_legacy', or "
            f"'bottom_padding'; got {mode!r}")
    if not pad:
        return Nx, Ny, Nz, (0, 0)
    if mode in ("mudpack", "mudpack_legacy"):
        g = fortran_helper(Nx, Ny, Nz, symmetric_xy=(mode == "mudpack"))
        return g.nx, g.ny, g.nz, (g.bfx, g.bfy)
    p = (Nx + Ny) // 6
    cx, cy = Nx + 2 * p, Ny + 2 * p
    cz = min(cx, cy)
    if cz < Nz:
        raise ValueError("bottom_padding calculation height cannot contain the input z domain: "
                         f"input_Nz={Nz}, calculated Nz={cz}")
    return cx, cy, cz, (p, p)


def centred_embedding(ny: int, nx: int, ty: int, tx: int):
    """Source and destination windows of a centred pad-or-crop (:1593-1609)."""
    sy0, sx0 = max(0, (ny - ty) // 2), max(0, (nx - tx) // 2)
    sy1, sx1 = sy0 + min(ny, ty), sx0 + min(nx, tx)
    dy0, dx0 = max(0, (ty - ny) // 2), max(0, (tx - nx) // 2)
    return (sy0, sy1, sx0, sx1), (dy0, dy0 + (sy1 - sy0), dx0, dx0 + (sx1 - sx0))


def sigma_window(nz_faces: int, ny: int, nx: int, observed_zyx, fov_start_yx):
    """Half-open (z0, z1, y0, y1, x0, x1) on the face grid (:2656-2676)."""
    nzo, nyo, nxo = observed_zyx
    bfy, bfx = fov_start_yx
    bfz = max(0, nz_faces - int(nzo))
    ax, ay, az = nxo // 8 + bfx, nyo // 8 + bfy, nzo // 4 + bfz
    z1 = max(1, nz_faces - 1 - az + 1)
    y1 = max(ay + 1, ny - 1 - ay + 1)
    x1 = max(ax + 1, nx - 1 - ax + 1)
    # a Python slice clips at the array end
    return 1, min(z1, nz_faces), ay, min(y1, ny), ax, min(x1, nx)
