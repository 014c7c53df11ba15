#!/usr/bin/env python3
"""Write golden input/output vectors for the NFGS plain iterate.

This script is the only place that imports the reference solver
(``nfgs_solver.py`` from the mounted reference tree).  It runs the
reference on the CPU in FP64 (and one FP32 case) at tiny grids and stores
inputs plus outputs as compressed ``.npz`` fixtures beside this file.  The
fixtures are data: seeded boundary maps, setup tensors, the gs_cfcf state,
two ``iteration_step`` states and a short ``iterate()`` trajectory.

Usage (in a container where the reference tree is mounted):

    python tests/golden/make_golden.py /root/reference

Nothing on a GPU machine runs this script; tests only read the ``.npz``.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def boundary(nyo: int, nxo: int, seed: int):
    """Two opposite Gaussian poles plus a smooth sheared horizontal field."""
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.arange(nyo, dtype=np.float64),
                       np.arange(nxo, dtype=np.float64), indexing="ij")
    cx, cy = 0.5 * (nxo - 1), 0.5 * (nyo - 1)
    th = rng.uniform(0.0, np.pi)
    sep = 0.3 * min(nxo, nyo)
    sig = 0.15 * min(nxo, nyo)
    dxp, dyp = 0.5 * sep * np.cos(th), 0.5 * sep * np.sin(th)
    g = lambda x0, y0: np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * sig * sig))
    bz = g(cx + dxp, cy + dyp) - g(cx - dxp, cy - dyp)
    bz += rng.normal(0.0, 1e-3, size=bz.shape)
    a1, a2 = rng.uniform(-0.3, 0.3, size=2)
    bx = a1 * np.sin(np.pi * y / (nyo - 1)) * bz + 0.05 * np.cos(0.4 * x)
    by = a2 * np.sin(np.pi * x / (nxo - 1)) * bz - 0.04 * np.sin(0.3 * y)
    return bx, by, bz


CASES = [
    # name, seed, observed (nxo, nyo, nz), rho, dtype, domain mode, pad, iterate kwargs
    ("mudpack_rho0", 1, (12, 10, 9), 0.0, "float64", "mudpack", True, {}),
    ("mudpack_rho05", 1, (12, 10, 9), 0.5, "float64", "mudpack", True, {}),
    ("mudpack_rho1", 1, (12, 10, 9), 1.0, "float64", "mudpack", True, {}),
    ("legacy_rho05", 2, (13, 11, 10), 0.5, "float64", "mudpack_legacy", True, {}),
    ("bottompad_rho05", 3, (10, 10, 8), 0.5, "float64", "bottom_padding", True, {}),
    ("unpadded_rho05", 4, (11, 9, 7), 0.5, "float64", "mudpack", False, {}),
    ("stop_rho05", 5, (12, 10, 9), 0.5, "float64", "mudpack", True,
     {"max_iter": 400, "patience": 5}),
    ("fp32_rho05", 1, (12, 10, 9), 0.5, "float32", "mudpack", True, {}),
    ("fp32_rho0", 2, (12, 10, 9), 0.0, "float32", "mudpack", True, {}),
    ("stall_rho1", 5, (12, 10, 9), 1.0, "float64", "mudpack", True,
     {"max_iter": 600, "patience": 4, "sigma_improvement_tol": 0.0}),
    ("tol_rho05", 1, (12, 10, 9), 0.5, "float64", "mudpack", True, {"tol": 0.85}),
    ("cube16_rho05", 6, (16, 16, 12), 0.5, "float64", "mudpack", True, {}),
]


def t2n(t):
    return t.detach().cpu().numpy().copy()


def run_case(ref, torch, name, seed, obs, rho, dtype_name, mode, pad, iter_kwargs):
    dtype = getattr(torch, dtype_name)
    nxo, nyo, nz = obs
    bx, by, bz = boundary(nyo, nxo, seed)
    kw = dict(
        device="cpu", dtype=dtype, lateral_wall="paper_bc1", pin_potential_boundaries=False,
        buffer_thickness=0, padded_buffer_zone=True, pad_calculation_domain=pad,
        calculation_domain_mode=mode, bottom_divergence_robin_weight=rho,
        mean_bz_policy="uniform_open", horizontal_zero_mode_policy="frozen_linear_harmonic",
        d0_zero_mode_policy="project", source_zero_mode_policy="project",
        harmonic_residual_policy="diagnose", poloidal_update_mode="direct",
        vertical_current_operator="staggered_face",
    )
    s = ref.NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    out = {
        "in_bx": bx, "in_by": by, "in_bz": bz,
        "calc_shape_zyx": np.array([s.Nz, s.Ny, s.Nx]),
        "fov_yx": np.array([s.fov_slice_yx[0].start, s.fov_slice_yx[0].stop,
                            s.fov_slice_yx[1].start, s.fov_slice_yx[1].stop]),
        "observed_zyx": np.array(s._observed_shape_zyx),
        "Bz_obs": t2n(s.Bz_obs), "Bx_obs": t2n(s.Bx_obs), "By_obs": t2n(s.By_obs),
        "Jz_obs": t2n(s.Jz_obs), "Az_bottom": t2n(s.Az_bottom),
        "Jz_T": t2n(s.Jz_toroidal_boundary), "B_background": t2n(s.B_background),
        "Phi_above_bottom": t2n(s.Phi_above_bottom), "Bz_above": t2n(s.Bz_above),
        "Bz_below": t2n(s.Bz_below), "D0": t2n(s.D0),
        "dphi_target": t2n(s._bottom_dphi_c_target), "H_k0": t2n(s._H_k0),
        "Bz_open_mean": np.array(float(s.Bz_open_mean)),
    }
    diag = s.gaussian_separation_diagnostics
    init = s.compute_initial_conditions(init_mode="gs_cfcf")
    for key, t in zip(("B", "J", "Phi_c", "Az"), init[:4]):
        out["init_" + key] = t2n(t)
    state = init
    for step in (1, 2):
        state = s.iteration_step(*state, mu=1.0)
        for key, t in zip(("B", "J", "Phi_c", "Az"), state[:4]):
            out[f"step{step}_" + key] = t2n(t)
        out[f"step{step}_sigma"] = np.array(s.compute_sigma(state[0], state[1]))
        out[f"step{step}_divB"] = np.array(s.compute_divB(state[0]))
        out[f"step{step}_energy"] = np.array(s.compute_energy(state[0]))
        out[f"step{step}_zero_mode_ratio"] = np.array(s._source_zero_mode_history[-1])
    out["init_sigma"] = np.array(s.compute_sigma(init[0], init[1]))
    kwargs = dict(init_mode="gs_cfcf", max_iter=30, tol=1e-5, mu=1, patience=40,
                  sigma_improvement_tol=1e-5, stop_mode="sigma_best", verbose=False)
    kwargs.update(iter_kwargs)
    hist = s.iterate(**kwargs)
    for key in ("sigma", "divB", "energy", "sigma_relative_change", "energy_relative_change"):
        out["hist_" + key] = np.array(hist[key], dtype=np.float64)
    scal = {k: hist[k] for k in ("best_sigma", "best_iter", "stop_reason", "final_iter",
                                 "final_sigma", "final_vs_best_B_relative",
                                 "last_sigma_improvement")}
    fields = s.fields_to_numpy(crop_to_observed_fov=True)
    for k, v in fields.items():
        out["out_" + k] = v
    sd = hist["source_diagnostics"]
    keep = ["bottom_reconstruction_rms", "correction_bottom_rms", "top_robin_rms",
            "top_phi_robin_rms", "top_upward_source_rms", "top_upward_source_ratio",
            "source_zero_mode_ratio_last", "source_zero_mode_ratio_max",
            "bottom_divergence_rms", "bottom_divergence_relative", "bottom_dphi_residual_rms",
            "Jz_projected_out_rms", "B_harm_over_Bobs_rms"]
    scal["source_diagnostics"] = {k: sd[k] for k in keep}
    scal["gaussian_diagnostics"] = {k: v for k, v in diag.items() if isinstance(v, (int, float))}
    scal["vertical_staggered_face"] = sd["vertical_staggered_face"]
    scal["toroidal"] = s.toroidal_boundary_diagnostics
    scal["case"] = dict(name=name, seed=seed, observed=obs, rho=rho, dtype=dtype_name, mode=mode, pad=pad,
                        iterate=kwargs)
    out["meta_json"] = np.array(json.dumps(scal))
    np.savez_compressed(HERE / f"{name}.npz", **out)
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
