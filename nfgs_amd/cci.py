"""Outward bottom-field comparison (the CCI measures), computed on the GPU.

The NFGS Gaussian separation splits a field's z = 0 plane into the part of
the potential field whose sources lie above the plane, B^> (the outward
Gaussian field, nfgs_solver.py:4921-4943).  Two reconstructions are compared
through their B^>: per pixel by |B^> - b^>| / |b^>| over the pixels where
|b^>| clears a threshold, and in L2 over the window.

Pipeline, all on the device:
  planes (3, ny, nx) --nfgs_setup_gaussian / nfgs_gaussian_above_field-->
  B^> (3, ny, nx) --nfgs_outward_errors--> raw window sums --> E^rel.

The public helpers keep the names and arguments of the reference's
(nfgs/_backend/scripts/plot_nfgs_td_robin_metrics_vs_er.py:240-326), so
scripts written against those carry over.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from . import _lib as L
from .solver import NFGSSolverGPU

# The split uses only the bottom plane; the constructor still needs a stack
# height (the reference's helper also builds a short one).
BGT_HELPER_NZ = 8

# Grid settings of a splitter: the plane is its own calculation grid (no
# padding, no Robin mixing), so B^> lives on exactly the caller's pixels.
_SPLIT_GRID = dict(bottom_divergence_projection_weight=0.0, bottom_divergence_robin_weight=0.0,
                   buffer_thickness=0, padded_buffer_zone=False, pad_calculation_domain=False,
                   lateral_wall="paper_bc1", pin_potential_boundaries=False)

# Separation policies a caller may choose (the evaluator's
# "gaussian_separation" block), with their defaults.
SPLIT_POLICIES = dict(mean_bz_policy="uniform_open",
                      horizontal_zero_mode_policy="frozen_linear_harmonic",
                      d0_zero_mode_policy="project", source_zero_mode_policy="project",
                      harmonic_residual_policy="diagnose", mean_mode_tolerance=1.0e-6,
                      source_zero_mode_tolerance=1.0e-2, gaussian_horizontal_operator="spectral")


def make_bgt_helper_solver(*, nx: int, ny: int, dx: float, dy: float, dz: float, device="cuda",
                           dtype=torch.float64, **policies) -> NFGSSolverGPU:
    """An unpadded NFGS instance whose only job is the bottom split
    (plot_nfgs_td_robin_metrics_vs_er.py:240-278).  `policies` override
    SPLIT_POLICIES; unknown names raise TypeError."""
    unknown = set(policies) - set(SPLIT_POLICIES)
    if unknown:
        raise TypeError(f"unknown separation policies: {sorted(unknown)}")
    chosen = {**SPLIT_POLICIES, **policies}
    return NFGSSolverGPU(int(nx), int(ny), BGT_HELPER_NZ, float(dx), float(dy), float(dz),
                         device=device, dtype=dtype, **_SPLIT_GRID, **chosen)


def outward_field(solver: NFGSSolverGPU, bx, by, bz) -> torch.Tensor:
    """B^> of one bottom plane triple as a (3, ny, nx) device tensor."""
    solver.set_bottom_vector_boundary(*(np.asarray(p, dtype=np.float64) for p in (bx, by, bz)),
                                      zero_net_current=False, smooth_passes=0, ijzbcor=3)
    return solver.gaussian_above_bottom_field()


def _bottom(fields):
    return tuple(fields[k][0] for k in ("Bx", "By", "Bz"))


def bottom_plane_bgt_nfgs(solver: NFGSSolverGPU, fields) -> np.ndarray:
    """B^> of the z = 0 plane of `fields` (Bx/By/Bz volumes), (3, ny, nx)
    FP64 NumPy (plot_nfgs_td_robin_metrics_vs_er.py:281-296)."""
    with torch.inference_mode():
        out = outward_field(solver, *_bottom(fields))
    return out.double().cpu().numpy()


@dataclass
class OutwardSums:
    """Window sums of nfgs_outward_errors for one target/candidate pair."""
    diff_sq: float          # sum |c - b|^2
    target_sq: float        # sum |b|^2
    target_max: float       # max |b|
    pixels: int             # pixels in the window
    kept: list              # per threshold: pixels with |b| above it
    rel_sum: list           # per threshold: sum of |c - b| / |b| over those

    def l2(self):
        """||c - b|| / ||b|| over the window (None for a zero target)."""
        return float(np.sqrt(self.diff_sq) / np.sqrt(self.target_sq)) if self.target_sq > 0 \
            else None

    def mean_rel(self, k: int):
        """Mean of |c - b| / |b| over the pixels above threshold k (None if none)."""
        return self.rel_sum[k] / self.kept[k] if self.kept[k] > 0 else None


def _plane_tensor(a) -> torch.Tensor:
    t = a if torch.is_tensor(a) else torch.as_tensor(np.asarray(a, dtype=np.float64))
    if t.dim() != 3 or int(t.shape[0]) != 3:
        raise ValueError(f"expected a (3, ny, nx) plane triple; got {tuple(t.shape)}")
    return t.to(device=torch.device("cuda", torch.cuda.current_device()),
                dtype=torch.float64).contiguous()


def outward_sums(target, candidate, window=(slice(None), slice(None)),
                 fractions: Sequence[float] = (), eps: float = 0.0) -> OutwardSums:
    """Run nfgs_outward_errors on two (3, ny, nx) plane triples (NumPy or
    device tensors).  Threshold k is fractions[k] * (max|b| + eps)."""
    if not torch.cuda.is_available():
        raise RuntimeError("nfgs_amd.cci needs a ROCm GPU: the HIP library has no CPU path")
    b, c = _plane_tensor(target), _plane_tensor(candidate)
    if b.shape != c.shape:
        raise ValueError(f"target and candidate shapes differ: {tuple(b.shape)} vs "
                         f"{tuple(c.shape)}")
    _, ny, nx = b.shape
    sy, sx = window
    y0, y1, ystep = sy.indices(ny)
    x0, x1, xstep = sx.indices(nx)
    if ystep != 1 or xstep != 1:
        raise ValueError("the comparison window must be unstrided")
    fr = [float(f) for f in fractions]
    arr = (C.c_double * max(len(fr), 1))(*fr)
    out = torch.zeros(4 + 2 * len(fr), device=b.device, dtype=torch.float64)
    lib = L.load()
    L.check(lib.nfgs_outward_errors(L.ptr(b), L.ptr(c), ny, nx, y0, y1, x0, x1, arr, len(fr),
                                    float(eps), L.ptr(out),
                                    torch.cuda.current_stream().cuda_stream), "outward errors")
    o = out.cpu().numpy()
    return OutwardSums(diff_sq=float(o[0]), target_sq=float(o[1]), target_max=float(o[2]),
                       pixels=int(o[3]), kept=[int(v) for v in o[4::2]],
                       rel_sum=[float(v) for v in o[5::2]])


def bottom_bgt_relative_error(solver: NFGSSolverGPU, ref_fields, sol_fields, *,
                              floor_frac: float = 1.0e-8, y_slice=None, x_slice=None) -> float:
    """E^rel_{B^>}: mean of |B^> - b^>| / |b^>| on z = 0 over the pixels with
    |b^>| > floor_frac * (max|b^>| + 1e-30), optionally on a y/x window
    (plot_nfgs_td_robin_metrics_vs_er.py:299-326)."""
    with torch.inference_mode():
        b = outward_field(solver, *_bottom(ref_fields)).clone()
        B = outward_field(solver, *_bottom(sol_fields))
        window = (slice(None) if y_slice is None else y_slice,
                  slice(None) if x_slice is None else x_slice)
        sums = outward_sums(b, B, window, (floor_frac,), eps=1.0e-30)
    value = sums.mean_rel(0)
    if value is None:
        raise RuntimeError("No valid bottom pixels for E^{rel}_{B^>} (|b^>| below floor).")
    return float(value)
