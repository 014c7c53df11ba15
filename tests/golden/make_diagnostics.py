#!/usr/bin/env python3
"""Write golden vectors of the per-iteration diagnostics and the fixed-point
stop (record_iteration_diagnostics=True, stop_mode="fixed_point").

Like make_golden.py this script imports the reference solver
(``nfgs_solver.py`` from the mounted reference tree) and runs it on the CPU
at tiny grids; the ``diag_*.npz`` fixtures it writes are
data (in ``diag/``): the seeded boundary maps, the diagnostics dict of two
``iteration_step`` calls from the gs_cfcf state, and the history of an
``iterate()`` run with every diagnostic list, its stop decisions and the
selected fields.

    python tests/golden/make_diagnostics.py /root/reference

Nothing on a GPU machine runs this script; tests only read the ``.npz``.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from make_golden import boundary, t2n  # noqa: E402

CASES = [
    # name, seed, observed (nxo, nyo, nz), rho, dtype, extra constructor kwargs, iterate kwargs
    ("diag_rho05", 1, (12, 10, 9), 0.5, "float64", {},
     {"max_iter": 30, "stop_mode": "sigma_best"}),
    ("diag_fp32_rho0", 2, (12, 10, 9), 0.0, "float32", {},
     {"max_iter": 30, "stop_mode": "sigma_best"}),
    # a weak-field floor and an alpha limit that bite, so both fractions are nonzero
    ("diag_floor_rho1", 3, (12, 10, 9), 1.0, "float64",
     {"weak_field_floor_b2": 2e-3, "alpha_limit_factor": 0.2},
     {"max_iter": 30, "stop_mode": "sigma_best"}),
    ("diag_fixed_point_rho05", 1, (12, 10, 9), 0.5, "float64", {},
     {"max_iter": 120, "stop_mode": "fixed_point"}),
]

DIAG_KEYS = ("alpha_clip_fraction", "az_map_residual_relative", "az_update_relative",
             "bcz_pde_residual_relative", "bottom_field_imprint_relative",
             "bottom_jz_imprint_relative", "h_k0_map_residual_relative", "h_k0_update_relative",
             "jz_pde_residual_relative", "mixed_row_residual_relative",
             "phi_map_residual_relative", "phi_update_relative",
             "poloidal_pde_residual_relative", "toroidal_pde_residual_relative",
             "weak_field_floor_fraction")


def run_case(ref, torch, name, seed, obs, rho, dtype_name, extra, iter_kwargs):
    dtype = getattr(torch, dtype_name)
    nxo, nyo, nz = obs
    bx, by, bz = boundary(nyo, nxo, seed)
    kw = dict(
        device="cpu", dtype=dtype, lateral_wall="paper_bc1", pin_potential_boundaries=False,
        buffer_thickness=0, padded_buffer_zone=True, pad_calculation_domain=True,
        calculation_domain_mode="mudpack", bottom_divergence_robin_weight=rho,
        mean_bz_policy="uniform_open", horizontal_zero_mode_policy="frozen_linear_harmonic",
        d0_zero_mode_policy="project", source_zero_mode_policy="project",
        harmonic_residual_policy="diagnose", poloidal_update_mode="direct",
        vertical_current_operator="staggered_face", record_iteration_diagnostics=True,
    )
    kw.update(extra)
    s = ref.NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    out = {"in_bx": bx, "in_by": by, "in_bz": bz}
    state = s.compute_initial_conditions(init_mode="gs_cfcf")
    steps = []
    for step in (1, 2):
        state = s.iteration_step(*state, mu=1.0)
        steps.append({k: float(s._last_iteration_diagnostics[k]) for k in DIAG_KEYS})
        for key, t in zip(("Phi_c", "Az"), state[2:4]):
            out[f"step{step}_" + key] = t2n(t)
    kwargs = dict(init_mode="gs_cfcf", tol=1e-5, mu=1, patience=40, sigma_improvement_tol=1e-5,
                  verbose=False)
    kwargs.update(iter_kwargs)
    hist = s.iterate(**kwargs)
    for key in ("sigma", "divB", "energy", "sigma_relative_change", "energy_relative_change",
                "fixed_point_criteria_met") + DIAG_KEYS:
        out["hist_" + key] = np.array(hist[key], dtype=np.float64)
    fields = s.fields_to_numpy(crop_to_observed_fov=True)
    for k, v in fields.items():
        out["out_" + k] = v
    scal = {k: hist[k] for k in ("best_sigma", "best_iter", "stop_reason", "final_iter",
                                 "stop_mode", "selected_state", "fixed_point_converged",
                                 "fixed_point_streak", "fixed_point_tolerances",
                                 "final_vs_best_B_relative", "iteration_diagnostics_definition")}
    scal["steps"] = steps
    scal["case"] = dict(name=name, seed=seed, observed=obs, rho=rho, dtype=dtype_name,
                        extra=extra, iterate=kwargs)
    out["meta_json"] = np.array(json.dumps(scal))
    (HERE / "diag").mkdir(exist_ok=True)
    np.savez_compressed(HERE / "diag" / f"{name}.npz", **out)
    print(name, s.Nz, s.Ny, s.Nx, hist["best_iter"], hist["final_iter"], hist["stop_reason"])


def main():
    only = set(sys.argv[2:])
    ref_root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    sys.path.insert(0, str(ref_root))
    import torch
    import nfgs_solver as ref  # noqa: E402  (reference solver, CPU only here)

    torch.set_num_threads(2)
    with torch.inference_mode():
        for case in CASES:
            if only and case[0] not in only:
                continue
            run_case(ref, torch, *case)


if __name__ == "__main__":
    main()
