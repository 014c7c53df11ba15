#!/usr/bin/env python3
"""Write golden vectors for relaxed and safeguarded iterations.

Runs the reference solver (``nfgs_solver.py`` from the mounted reference
tree) on the CPU with mu < 1 (relaxed scalars), adaptive mu with rollbacks,
and the polarity masks of the bottom parallel current, at tiny grids, and
stores inputs and outputs as compressed ``.npz`` fixtures in
``tests/golden/relax/``: the gs_cfcf state, two ``iteration_step`` states at
the case's mu and mask, a fixed-length ``iterate()`` history (sigma, divB,
energy, mu, rollback) and the cropped output fields.

Usage (in a container where the reference tree is mounted):

    python tests/golden/make_relax.py /root/reference

Nothing on a GPU machine runs this script; tests only read the ``.npz``.
"""
from __future__ import annotations

import json
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "relax"
sys.path.insert(0, str(HERE))
from make_golden import boundary, t2n  # noqa: E402  (the same seeded boundary maps)

CASES = [
    # name, seed, observed, rho, dtype, solver kwargs, step kwargs, iterate kwargs
    ("mu05", 1, (12, 10, 9), 0.5, "float64", {}, dict(mu=0.5),
     dict(mu=0.5, max_iter=30)),
    ("mu07_neg", 2, (12, 10, 9), 0.0, "float64", {}, dict(mu=0.7, polarity_mask="negative"),
     dict(mu=0.7, polarity_mask="negative", max_iter=30)),
    ("pos", 3, (12, 10, 9), 1.0, "float64", {}, dict(mu=1.0, polarity_mask="positive"),
     dict(mu=1.0, polarity_mask="positive", max_iter=30)),
    # mu grows 0.5 -> 1.0 over improving streaks, then six rollbacks stop it
    ("adaptive", 8, (12, 10, 9), 0.5, "float64", {}, dict(mu=0.5),
     dict(mu=0.5, adaptive_mu=True, divergence_factor=1.0, growth_streak=2, mu_growth=1.3,
          mu_backoff=0.6, max_rollbacks=6, max_iter=80)),
    # rollbacks with backoff to the floor, stopping after three
    ("adaptive_cap", 3, (12, 10, 9), 0.5, "float64", {}, dict(mu=0.8),
     dict(mu=0.8, adaptive_mu=True, divergence_factor=1.0, mu_backoff=0.2, mu_min=0.1,
          max_rollbacks=3, max_iter=120)),
    ("mu06_coupled", 6, (12, 10, 9), 0.5, "float64",
     dict(horizontal_zero_mode_policy="coupled_k0"), dict(mu=0.6), dict(mu=0.6, max_iter=30)),
    ("mu05_fp32", 1, (12, 10, 9), 0.5, "float32", {}, dict(mu=0.5),
     dict(mu=0.5, max_iter=30)),
]


def run_case(ref, torch, name, seed, obs, rho, dtype_name, extra, step_kw, iter_kw):
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
    kw.update(extra)
    s = ref.NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    out = {"in_bx": bx, "in_by": by, "in_bz": bz,
           "calc_shape_zyx": np.array([s.Nz, s.Ny, s.Nx])}
    state = s.compute_initial_conditions(init_mode="gs_cfcf")
    for key, t in zip(("B", "J", "Phi_c", "Az"), state[:4]):
        out["init_" + key] = t2n(t)
    for step in (1, 2):
        state = s.iteration_step(*state, **step_kw)
        for key, t in zip(("B", "J", "Phi_c", "Az"), state[:4]):
            out[f"step{step}_" + key] = t2n(t)
        out[f"step{step}_sigma"] = np.array(s.compute_sigma(state[0], state[1]))
    s = ref.NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    kwargs = dict(init_mode="gs_cfcf", tol=1e-5, patience=40, sigma_improvement_tol=0.0,
                  stop_mode="sigma_best", verbose=False)
    kwargs.update(iter_kw)
    hist = s.iterate(**kwargs)
    for key in ("sigma", "divB", "energy", "mu", "rollback"):
        out["hist_" + key] = np.array(hist[key], dtype=np.float64)
    fields = s.fields_to_numpy(crop_to_observed_fov=True)
    for k, v in fields.items():
        out["out_" + k] = v
    meta = {"case": dict(name=name, seed=seed, observed=obs, rho=rho, dtype=dtype_name,
                         extra=extra),
            "step": step_kw, "iterate": kwargs,
            "hist": {k: hist[k] for k in ("best_sigma", "best_iter", "stop_reason", "final_iter",
                                          "final_sigma", "rollbacks")}}
    meta["hist"] = {k: (None if isinstance(v, float) and not math.isfinite(v) else v)
                    for k, v in meta["hist"].items()}
    out["meta_json"] = np.array(json.dumps(meta))
    OUT.mkdir(exist_ok=True)
    np.savez_compressed(OUT / f"{name}.npz", **out)
    print(name, s.Nz, s.Ny, s.Nx, hist["best_iter"], hist["stop_reason"], hist["rollbacks"],
          len(hist["sigma"]), sum(hist["rollback"]))


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
