#!/usr/bin/env python3
"""Write golden vectors for the coupled horizontal-harmonic closures.

Runs the reference solver (``nfgs_solver.py`` from the mounted reference
tree) on the CPU with ``horizontal_zero_mode_policy="coupled_k0"`` and
``"constrained_spline_harmonic"`` at tiny grids and stores inputs and
outputs as compressed ``.npz`` fixtures in ``tests/golden/k0/``: the setup's
bottom harmonic target, the gs_cfcf state, two ``iteration_step`` states with
the harmonic profile after each, a fixed-length ``iterate()`` trajectory, the
final profile, the cropped output fields and the closure diagnostics.

Usage (in a container where the reference tree is mounted):

    python tests/golden/make_k0.py /root/reference

Nothing on a GPU machine runs this script; tests only read the ``.npz``.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "k0"
sys.path.insert(0, str(HERE))
from make_golden import boundary, t2n  # noqa: E402  (the same seeded boundary maps)

CASES = [
    # name, seed, observed (nxo, nyo, nz), rho, dtype, closure kwargs
    ("coupled_rho05", 1, (12, 10, 9), 0.5, "float64",
     dict(horizontal_zero_mode_policy="coupled_k0")),
    ("coupled_rho0", 2, (12, 10, 9), 0.0, "float64",
     dict(horizontal_zero_mode_policy="coupled_k0", k0_ode_smoothness=1e-2)),
    ("coupled_legacy_rho05", 3, (12, 10, 9), 0.5, "float64",
     dict(horizontal_zero_mode_policy="coupled_k0",
          coupled_k0_discretization="legacy_plane_centred")),
    ("coupled_mu_rho1", 4, (12, 10, 9), 1.0, "float64",
     dict(horizontal_zero_mode_policy="coupled_k0", k0_update_mu=0.5)),
    ("spline_rho05", 1, (12, 10, 9), 0.5, "float64",
     dict(horizontal_zero_mode_policy="constrained_spline_harmonic")),
    ("spline_basis6_rho0", 5, (12, 10, 9), 0.0, "float64",
     dict(horizontal_zero_mode_policy="constrained_spline_harmonic", k0_spline_basis_size=6,
          k0_ode_smoothness=0.0)),
    ("spline_cube16", 6, (16, 16, 12), 0.5, "float64",
     dict(horizontal_zero_mode_policy="constrained_spline_harmonic")),
    ("coupled_fp32", 1, (12, 10, 9), 0.5, "float32",
     dict(horizontal_zero_mode_policy="coupled_k0")),
    ("spline_fp32", 1, (12, 10, 9), 0.5, "float32",
     dict(horizontal_zero_mode_policy="constrained_spline_harmonic")),
]

KEEP_K0 = ("ode_residual_rms", "ode_residual_relative", "ode_residual_relative_max",
           "face_ode_residual_rms", "face_ode_residual_relative", "gridscale_profile_rms",
           "gridscale_profile_relative", "top_derivative", "profile_max_abs", "profile_top",
           "desired_mean_current_rms")


def k0_diag(s):
    d = s._k0_diagnostics or {}
    return {k: d[k] for k in KEEP_K0 if k in d}


def run_case(ref, torch, name, seed, obs, rho, dtype_name, closure):
    dtype = getattr(torch, dtype_name)
    nxo, nyo, nz = obs
    bx, by, bz = boundary(nyo, nxo, seed)
    kw = dict(
        device="cpu", dtype=dtype, lateral_wall="paper_bc1", pin_potential_boundaries=False,
        buffer_thickness=0, padded_buffer_zone=True, pad_calculation_domain=True,
        calculation_domain_mode="mudpack", bottom_divergence_robin_weight=rho,
        mean_bz_policy="uniform_open", d0_zero_mode_policy="project",
        source_zero_mode_policy="project", harmonic_residual_policy="diagnose",
        poloidal_update_mode="direct", vertical_current_operator="staggered_face",
    )
    kw.update(closure)
    s = ref.NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    out = {
        "in_bx": bx, "in_by": by, "in_bz": bz,
        "calc_shape_zyx": np.array([s.Nz, s.Ny, s.Nx]),
        "H_bottom_target": t2n(s._H_k0_bottom_target), "H_k0_setup": t2n(s._H_k0),
    }
    meta = {"case": dict(name=name, seed=seed, observed=obs, rho=rho, dtype=dtype_name,
                         closure=closure), "setup_k0": k0_diag(s)}
    init = s.compute_initial_conditions(init_mode="gs_cfcf")
    for key, t in zip(("B", "J", "Phi_c", "Az"), init[:4]):
        out["init_" + key] = t2n(t)
    out["init_H_k0"] = t2n(s._H_k0)
    state = init
    for step in (1, 2):
        state = s.iteration_step(*state, mu=1.0)
        for key, t in zip(("B", "J", "Phi_c", "Az"), state[:4]):
            out[f"step{step}_" + key] = t2n(t)
        out[f"step{step}_H_k0"] = t2n(s._H_k0)
        out[f"step{step}_sigma"] = np.array(s.compute_sigma(state[0], state[1]))
        meta[f"step{step}_k0"] = k0_diag(s)
    # a fresh solver for the trajectory (the steps above changed its profile)
    s = ref.NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    kwargs = dict(init_mode="gs_cfcf", max_iter=30, tol=1e-5, mu=1, patience=40,
                  sigma_improvement_tol=0.0, stop_mode="sigma_best", verbose=False)
    hist = s.iterate(**kwargs)
    for key in ("sigma", "divB", "energy"):
        out["hist_" + key] = np.array(hist[key], dtype=np.float64)
    out["final_H_k0"] = t2n(s._H_k0)
    fields = s.fields_to_numpy(crop_to_observed_fov=True)
    for k, v in fields.items():
        out["out_" + k] = v
    meta["hist"] = {k: hist[k] for k in ("best_sigma", "best_iter", "stop_reason", "final_iter",
                                         "final_sigma")}
    meta["final_k0"] = k0_diag(s)
    meta["init_diagnostics_k0"] = {k: v for k, v in
                                   (hist["init_diagnostics"].get("coupled_k0") or {}).items()
                                   if isinstance(v, (int, float, str, bool, list))}
    meta["iterate"] = kwargs
    out["meta_json"] = np.array(json.dumps(meta))
    OUT.mkdir(exist_ok=True)
    np.savez_compressed(OUT / f"{name}.npz", **out)
    print(name, s.Nz, s.Ny, s.Nx, hist["best_iter"], hist["stop_reason"], hist["best_sigma"])


def main():
    only = set(sys.argv[2:])
    ref_root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    sys.path.insert(0, str(ref_root))
    import torch
    import nfgs_solver as ref  # noqa: E402  (reference solver, CPU only here)

    torch.set_num_threads(4)
    with torch.inference_mode():
        for case in CASES:
            if only and case[0] not in only:
                continue
            run_case(ref, torch, *case)


if __name__ == "__main__":
    main()
