"""Reconstruct a vector photospheric boundary stored as Bx/By/Bz, dx/dy/dz and Nz.

The on-disk formats of the reference's examples/reconstruct_npz.py:1-67:
the input NPZ holds the observed Bx, By, Bz planes (ny, nx), the spacings dx,
dy, dz and the height Nz; the output NPZ holds the six fields of
fields_to_numpy(crop_to_observed_fov=True) (Bx, By, Bz, Jx, Jy, Jz, each
(Nz, ny, nx)), and a JSON summary beside it (same stem, .json) holds rho,
best_iter, best_sigma and stop_reason.  FP32 solve, gs_cfcf start, the
sigma_best stop rule with the reference example's settings.

    python -m nfgs_amd.reconstruct_npz boundary.npz --output out.npz --rho 0.5
"""
from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np
import torch

from .solver import NFGSSolverGPU


def read_boundary(path):
    """(bx, by, bz, nz, (dx, dy, dz)) of an input NPZ; nothing in it is unpickled."""
    with np.load(path, allow_pickle=False) as d:
        bx, by, bz = (np.asarray(d[k]) for k in ("Bx", "By", "Bz"))
        nz = int(d["Nz"])
        spacing = tuple(float(d[k]) for k in ("dx", "dy", "dz"))
    if bz.ndim != 2 or bx.shape != bz.shape or by.shape != bz.shape:
        raise ValueError(f"{path}: Bx, By, Bz must be (ny, nx) planes of one shape")
    return bx, by, bz, nz, spacing


def reconstruct(bx, by, bz, nz, spacing, *, rho=0.5, max_iter=3000, device="cuda",
                verbose=True):
    """Solve one boundary as the reference example does; returns (solver, history)."""
    ny, nx = bz.shape
    solver = NFGSSolverGPU(nx, ny, nz, *spacing, device=device, dtype=torch.float32,
                           calculation_domain_mode="mudpack",
                           bottom_divergence_robin_weight=rho,
                           horizontal_zero_mode_policy="frozen_linear_harmonic",
                           vertical_current_operator="staggered_face")
    solver.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0,
                                      ijzbcor=3)
    with torch.inference_mode():
        history = solver.iterate(init_mode="gs_cfcf", max_iter=max_iter, tol=1e-5, mu=1,
                                 patience=40, sigma_improvement_tol=1e-5, stop_mode="sigma_best",
                                 verbose=verbose)
    return solver, history


def write_outputs(solver, history, rho, output: Path):
    """The output NPZ (six fields) and its JSON summary."""
    output.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(output, **solver.fields_to_numpy(crop_to_observed_fov=True))
    summary = {"rho": rho, "best_iter": history["best_iter"], "best_sigma": history["best_sigma"],
               "stop_reason": history["stop_reason"]}
    output.with_suffix(".json").write_text(json.dumps(summary, indent=2) + "\n")
    return summary


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("boundary", type=Path)
    p.add_argument("--output", type=Path, default=Path("reconstruction.npz"))
    p.add_argument("--rho", type=float, default=0.5)
    p.add_argument("--device", default="auto")
    p.add_argument("--max-iter", type=int, default=3000)
    a = p.parse_args(argv)
    if not 0 <= a.rho <= 1:
        p.error("--rho must be in [0,1]")
    device = "cuda" if a.device == "auto" else a.device
    if not str(device).startswith("cuda"):
        p.error("nfgs_amd solves on a ROCm GPU only (--device cuda or auto)")
    bx, by, bz, nz, spacing = read_boundary(a.boundary)
    solver, history = reconstruct(bx, by, bz, nz, spacing, rho=a.rho, max_iter=a.max_iter,
                                  device=device)
    write_outputs(solver, history, a.rho, a.output)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
