"""Manufactured solutions of the reference's verification harnesses, in NumPy.

The reference checks its per-stage operators against semidiscrete
manufactured fields (nfgs/_backend/experiments/nfgs_paper/
modal_verification/run.py:131-201 and extended_manufactured/run.py:163-446):
cosine/sine modes of the lateral finite-difference eigenbasis times analytic
z profiles, so the lateral operator is exact and the errors measure the
second-order z discretisation and the boundary rows.  The same fields are
built here from the same formulas for the GPU and oracle tests.
"""
from __future__ import annotations

import math

import numpy as np

# (mx, my, amplitude, a) of f(z) = exp(-k z) + a z^2 (modal_verification/run.py:27-31)
MODAL_CASES = {
    "homogeneous": ((1, 0, 1.0, 0.0),),
    "forced": ((1, 0, 1.0, 0.20),),
    "multimode": ((1, 1, 1.0, 0.12), (2, 1, 0.30, -0.05)),
}
# extended_manufactured/run.py:40-51
MIXED_COMPONENTS = ((1, 1, 1.0, 0.10), (2, 1, -0.25, -0.04))
PHI_COMPONENTS = ((1, 1, 0.70, 0.08), (2, 1, -0.20, -0.035))
AZ_COMPONENTS = ((1, 1, 0.45, 0.35), (2, 1, -0.18, -0.20))


def neumann_lambda(nx, ny, dx, dy):
    """Endpoint-Neumann eigenvalues (ny, nx), nfgs_solver.py:1098-1109."""
    kx, ky = np.arange(nx), np.arange(ny)
    lx = -4.0 * np.sin(np.pi * kx / (2.0 * max(nx - 1, 1))) ** 2 / dx ** 2
    ly = -4.0 * np.sin(np.pi * ky / (2.0 * max(ny - 1, 1))) ** 2 / dy ** 2
    return ly[:, None] + lx[None, :]


def dirichlet_lambda(nx, ny, dx, dy):
    """Endpoint-Dirichlet interior eigenvalues (ny-2, nx-2), :1083-1096."""
    nxi, nyi = nx - 2, ny - 2
    kx, ky = np.arange(1, nxi + 1), np.arange(1, nyi + 1)
    lx = -4.0 * np.sin(np.pi * kx / (2.0 * (nxi + 1))) ** 2 / dx ** 2
    ly = -4.0 * np.sin(np.pi * ky / (2.0 * (nyi + 1))) ** 2 / dy ** 2
    return ly[:, None] + lx[None, :]


def _axes(nx, ny, nz, dx, dy, dz):
    return np.arange(nx) * dx, np.arange(ny) * dy, np.arange(nz) * dz


def cosine_mode(nx, ny, dx, dy, mx, my):
    x, y, _ = _axes(nx, ny, 1, dx, dy, 1.0)
    lx = max((nx - 1) * dx, dx)
    ly = max((ny - 1) * dy, dy)
    return np.cos(math.pi * my * y / ly)[:, None] * np.cos(math.pi * mx * x / lx)[None, :]


def modal_fields(nx, ny, nz, dx, dy, dz, components):
    """Phi, its source, B_z = lap_xy Phi, the bottom values and the top Robin
    values g = (d_z + k) Phi of modal_verification/run.py:144-201."""
    _, _, z = _axes(nx, ny, nz, dx, dy, dz)
    height = (nz - 1) * dz
    lam_xy = neumann_lambda(nx, ny, dx, dy)
    phi = np.zeros((nz, ny, nx))
    source, bz = np.zeros_like(phi), np.zeros_like(phi)
    bottom, top = np.zeros((ny, nx)), np.zeros((ny, nx))
    for mx, my, amp, a in components:
        lam = float(lam_xy[my, mx])
        kh = math.sqrt(max(-lam, 0.0))
        mode = cosine_mode(nx, ny, dx, dy, mx, my)
        f = np.exp(-kh * z) + a * z * z
        src = 2.0 * a - kh * kh * a * z * z
        g_top = 2.0 * a * height + kh * a * height * height
        phi += amp * f[:, None, None] * mode
        source += amp * src[:, None, None] * mode
        bz += amp * lam * f[:, None, None] * mode
        bottom += amp * mode
        top += amp * g_top * mode
    return dict(phi=phi, source=source, bz=bz, bottom=bottom, top=top)


def bottom_derivative(field, dz):
    """Second-order one-sided d/dz at z = 0 (extended_manufactured/run.py:189-192)."""
    return (-3.0 * field[0] + 4.0 * field[1] - field[2]) / (2.0 * dz)


def complete_fields(nx, ny, nz, dx, dy, dz):
    """Phi_c and Az with their sources and the analytic B_c and J of the
    complete PT problem (extended_manufactured/run.py:336-446)."""
    x, y, z = _axes(nx, ny, nz, dx, dy, dz)
    lx, ly, height = (nx - 1) * dx, (ny - 1) * dy, (nz - 1) * dz
    lam_n = neumann_lambda(nx, ny, dx, dy)
    lam_d = dirichlet_lambda(nx, ny, dx, dy)
    shape = (nz, ny, nx)
    phi, source_bcz = np.zeros(shape), np.zeros(shape)
    az, source_jz = np.zeros(shape), np.zeros(shape)
    b_ref, j_ref = np.zeros((3,) + shape), np.zeros((3,) + shape)
    for mx, my, amp, a in PHI_COMPONENTS:
        kx, ky = math.pi * mx / lx, math.pi * my / ly
        cx, sx = np.cos(kx * x), np.sin(kx * x)
        cy, sy = np.cos(ky * y), np.sin(ky * y)
        coscos, sincos, cossin = cy[:, None] * cx[None], cy[:, None] * sx[None], \
            sy[:, None] * cx[None]
        lam = float(lam_n[my, mx])
        kh = math.sqrt(max(-lam, 0.0))
        dist = height - z
        f = np.exp(-kh * z) + a * dist ** 4
        fp = -kh * np.exp(-kh * z) - 4.0 * a * dist ** 3
        fpp = kh * kh * np.exp(-kh * z) + 12.0 * a * dist ** 2
        ss = fpp + lam * f
        phi += amp * f[:, None, None] * coscos
        source_bcz += amp * lam * ss[:, None, None] * coscos
        b_ref[0] += amp * kx * fp[:, None, None] * sincos
        b_ref[1] += amp * ky * fp[:, None, None] * cossin
        b_ref[2] += amp * lam * f[:, None, None] * coscos
        j_ref[0] += -amp * ky * ss[:, None, None] * cossin
        j_ref[1] += amp * kx * ss[:, None, None] * sincos
    zs = z / height
    for mx, my, amp, c in AZ_COMPONENTS:
        kx, ky = math.pi * mx / lx, math.pi * my / ly
        cx, sx = np.cos(kx * x), np.sin(kx * x)
        cy, sy = np.cos(ky * y), np.sin(ky * y)
        sinsin, sincos, cossin = sy[:, None] * sx[None], cy[:, None] * sx[None], \
            sy[:, None] * cx[None]
        lam = float(lam_d[my - 1, mx - 1])
        e = np.exp(c * zs)
        g = (1.0 - zs) * e
        gp = e * (c * (1.0 - zs) - 1.0) / height
        gpp = e * (c * c * (1.0 - zs) - 2.0 * c) / (height * height)
        rhs = gpp + lam * g
        az += amp * g[:, None, None] * sinsin
        source_jz += -amp * lam * rhs[:, None, None] * sinsin
        b_ref[0] += amp * ky * g[:, None, None] * sincos
        b_ref[1] += -amp * kx * g[:, None, None] * cossin
        j_ref[0] += amp * kx * gp[:, None, None] * cossin
        j_ref[1] += amp * ky * gp[:, None, None] * sincos
        j_ref[2] += -amp * lam * g[:, None, None] * sinsin
    return dict(phi=phi, source_bcz=source_bcz, az=az, source_jz=source_jz, b_ref=b_ref,
                j_ref=j_ref)


def observed_order(errors, hs):
    """Observed orders between successive grids (coarse to fine)."""
    return [math.log(e0 / e1) / math.log(h0 / h1)
            for (e0, h0), (e1, h1) in zip(zip(errors, hs), zip(errors[1:], hs[1:]))]
