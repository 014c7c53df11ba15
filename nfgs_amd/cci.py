"""Bottom-plane re-separation of output fields (the CCI comparison).

E^rel_{B^>} = mean over the bottom plane of |b^> - B^>| / |b^>|, where b^> and
B^> are the outward Gaussian parts (NFGSSolverGPU.gaussian_above_bottom_field,
nfgs_solver.py:4921-4943) of a reference field's and a solution's z = 0
planes, pixels with |b^>| at or below floor_frac * max |b^>| left out.  Same
names, arguments and results as the reference's helpers in
nfgs/_backend/scripts/plot_nfgs_td_robin_metrics_vs_er.py:240-326.  The
separation runs on the GPU (nfgs_setup_gaussian, nfgs_gaussian_above_field);
the plane-sized mean is host arithmetic, as in the reference.
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import NFGSSolverGPU

# The bottom split does not use Nz; the helper keeps a small stack for the
# constructor (plot_nfgs_td_robin_metrics_vs_er.py:100).
BGT_HELPER_NZ = 8


def make_bgt_helper_solver(*, nx: int, ny: int, dx: float, dy: float, dz: float, device="cuda",
                           dtype=torch.float64, mean_bz_policy="uniform_open",
                           horizontal_zero_mode_policy="frozen_linear_harmonic",
                           d0_zero_mode_policy="project", source_zero_mode_policy="project",
                           harmonic_residual_policy="diagnose", mean_mode_tolerance=1.0e-6,
                           source_zero_mode_tolerance=1.0e-2,
                           gaussian_horizontal_operator="spectral") -> NFGSSolverGPU:
    """An unpadded NFGS instance used only for the bottom split
    (plot_nfgs_td_robin_metrics_vs_er.py:240-278; the policies are keywords
    here instead of an argparse namespace)."""
    return NFGSSolverGPU(
        int(nx), int(ny), BGT_HELPER_NZ, float(dx), float(dy), float(dz), device=device,
        dtype=dtype, mean_bz_policy=mean_bz_policy,
        horizontal_zero_mode_policy=horizontal_zero_mode_policy,
        d0_zero_mode_policy=d0_zero_mode_policy, source_zero_mode_policy=source_zero_mode_policy,
        harmonic_residual_policy=harmonic_residual_policy, mean_mode_tolerance=mean_mode_tolerance,
        source_zero_mode_tolerance=source_zero_mode_tolerance,
        bottom_divergence_projection_weight=0.0, bottom_divergence_robin_weight=0.0,
        buffer_thickness=0, padded_buffer_zone=False, pad_calculation_domain=False,
        lateral_wall="paper_bc1", pin_potential_boundaries=False,
        gaussian_horizontal_operator=gaussian_horizontal_operator)


def bottom_plane_bgt_nfgs(solver: NFGSSolverGPU, fields) -> np.ndarray:
    """B^> of fields' z = 0 plane, (3, ny, nx) FP64 NumPy
    (plot_nfgs_td_robin_metrics_vs_er.py:281-296)."""
    solver.set_bottom_vector_boundary(
        np.asarray(fields["Bx"][0], dtype=np.float64),
        np.asarray(fields["By"][0], dtype=np.float64),
        np.asarray(fields["Bz"][0], dtype=np.float64),
        zero_net_current=False, smooth_passes=0, ijzbcor=3)
    with torch.inference_mode():
        bgt = solver.gaussian_above_bottom_field()
    return np.asarray(bgt.detach().cpu(), dtype=np.float64)


def bottom_bgt_relative_error(solver: NFGSSolverGPU, ref_fields, sol_fields, *,
                              floor_frac: float = 1.0e-8, y_slice=None, x_slice=None) -> float:
    """E^rel_{B^>} on z = 0, optionally on a y/x window
    (plot_nfgs_td_robin_metrics_vs_er.py:299-326)."""
    b = bottom_plane_bgt_nfgs(solver, ref_fields)
    B = bottom_plane_bgt_nfgs(solver, sol_fields)
    sy = slice(None) if y_slice is None else y_slice
    sx = slice(None) if x_slice is None else x_slice
    b = b[:, sy, sx]
    B = B[:, sy, sx]
    b_mag = np.linalg.norm(b, axis=0)
    floor = float(floor_frac) * float(np.max(b_mag) + 1.0e-30)
    mask = b_mag > floor
    if not np.any(mask):
        raise RuntimeError("No valid bottom pixels for E^{rel}_{B^>} (|b^>| below floor).")
    diff = np.linalg.norm(b - B, axis=0)
    return float(np.mean(diff[mask] / b_mag[mask]))
