#!/usr/bin/env python3
"""Write golden vectors for the soft_cfcf initializer.

Runs the reference solver (``nfgs_solver.py`` from the mounted reference
tree) on the CPU with ``init_mode="soft_cfcf"`` at tiny grids and stores
inputs and outputs as compressed ``.npz`` fixtures in ``tests/golden/soft/``:
the modal bottom data and wavenumbers, the modal profiles of both scalars,
the initial state (B, J, Phi_c, Az), its soft_cfcf diagnostics, a
fixed-length ``iterate()`` trajectory from it and the cropped output fields.

Usage (in a container where the reference tree is mounted):

    python tests/golden/make_soft.py /root/reference

Nothing on a GPU machine runs this script; tests only read the ``.npz``.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
OUT = HERE / "soft"
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))
from make_golden import boundary, t2n  # noqa: E402  (the same seeded boundary maps)

CASES = [
    # name, seed, observed (nxo, nyo, nz), rho, dtype, extra solver kwargs
    ("soft_rho05", 1, (12, 10, 9), 0.5, "float64", {}),
    ("soft_rho0", 2, (12, 10, 9), 0.0, "float64", {}),
    ("soft_eps_ridge", 3, (12, 10, 9), 1.0, "float64",
     dict(soft_cfcf_epsilon=0.3, soft_cfcf_normal_ridge=1e-9)),
    ("soft_eps0", 4, (12, 10, 9), 0.5, "float64", dict(soft_cfcf_epsilon=0.0)),
    ("soft_cube16", 6, (16, 16, 16), 0.5, "float64", {}),
    ("soft_fp32", 1, (12, 10, 9), 0.5, "float32", {}),
    ("soft_coupled", 5, (12, 10, 9), 0.5, "float64",
     dict(horizontal_zero_mode_policy="coupled_k0")),
]


def run_case(ref, torch, name, seed, obs, rho, dtype_name, extra):
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
    # the modal inputs and the modal profiles of both scalars
    phi_bottom_hat = s._dct2(s.Phi_above_bottom)
    phi_bottom_hat[0, 0] = 0.0
    az_bottom_hat = s._dst1(s._dst1(s.Az_bottom[1:-1, 1:-1], -1), -2)
    az_k = torch.sqrt(torch.clamp(-s._dirichlet_lam_xy_interior(), min=0.0))
    out["phi_k"] = t2n(s._nfgs_k)
    out["phi_bottom_hat"] = t2n(phi_bottom_hat)
    out["az_k"] = t2n(az_k)
    out["az_bottom_hat"] = t2n(az_bottom_hat)
    out["phi_hat"] = t2n(s._solve_soft_cfcf_modes(s._nfgs_k, phi_bottom_hat, kind="phi"))
    out["az_hat"] = t2n(s._solve_soft_cfcf_modes(az_k, az_bottom_hat, kind="az"))
    init = s.compute_initial_conditions(init_mode="soft_cfcf")
    for key, t in zip(("B", "J", "Phi_c", "Az", "B_c"), init):
        out["init_" + key] = t2n(t)
    out["init_sigma"] = np.array(s.compute_sigma(init[0], init[1]))
    meta = {"case": dict(name=name, seed=seed, observed=obs, rho=rho, dtype=dtype_name,
                         extra=extra),
            "soft_cfcf_epsilon": s.soft_cfcf_epsilon,
            "init_diagnostics": {k: v for k, v in s._init_diagnostics.items()
                                 if isinstance(v, (int, float, str))},
            "soft": dict(s._init_diagnostics["soft_cfcf"])}
    s = ref.NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    kwargs = dict(init_mode="soft_cfcf", max_iter=30, tol=1e-5, mu=1, patience=40,
                  sigma_improvement_tol=0.0, stop_mode="sigma_best", verbose=False)
    hist = s.iterate(**kwargs)
    for key in ("sigma", "divB", "energy"):
        out["hist_" + key] = np.array(hist[key], dtype=np.float64)
    fields = s.fields_to_numpy(crop_to_observed_fov=True)
    for k, v in fields.items():
        out["out_" + k] = v
    meta["hist"] = {k: hist[k] for k in ("best_sigma", "best_iter", "stop_reason", "final_iter",
                                         "final_sigma")}
    meta["iterate"] = kwargs
    out["meta_json"] = np.array(json.dumps(meta))
    OUT.mkdir(exist_ok=True)
    np.savez_compressed(OUT / f"{name}.npz", **out)
    print(name, s.Nz, s.Ny, s.Nx, hist["best_iter"], hist["stop_reason"], hist["best_sigma"])


def run_digest_case(ref, torch, dtype_name):
    """The start at a 64^3 observed multipole boundary (66 x 66 x 65
    calculation grid), held as digests of its volumes (tests/digest.py)."""
    from digest import digest, pack  # noqa: E402
    from nfgs_amd.boundary import multipole_boundary  # noqa: E402  (seeded input maps)

    dtype = getattr(torch, dtype_name)
    np_dt = np.float64 if dtype_name == "float64" else np.float32
    bx, by, bz, _ = multipole_boundary(64, 64, seed=7, sigma=(4.0, 10.0))
    bx, by, bz = (a.astype(np_dt) for a in (bx, by, bz))
    kw = dict(
        device="cpu", dtype=dtype, lateral_wall="paper_bc1", pin_potential_boundaries=False,
        buffer_thickness=0, padded_buffer_zone=True, pad_calculation_domain=True,
        calculation_domain_mode="mudpack", bottom_divergence_robin_weight=0.5,
        mean_bz_policy="uniform_open", d0_zero_mode_policy="project",
        source_zero_mode_policy="project", harmonic_residual_policy="diagnose",
        poloidal_update_mode="direct", vertical_current_operator="staggered_face",
    )
    s = ref.NFGSSolverGPU(64, 64, 64, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    B, J, phi, az, _ = s.compute_initial_conditions(init_mode="soft_cfcf")
    out = {"in_bx": bx, "in_by": by, "in_bz": bz}
    for key, t in (("B", B), ("J", J), ("Phi_c", phi), ("Az", az)):
        out.update(pack(key, digest(t2n(t))))
    meta = {"case": dict(name=f"soft64_{dtype_name}", observed=[64, 64, 64], rho=0.5,
                         dtype=dtype_name, extra={}),
            "soft": dict(s._init_diagnostics["soft_cfcf"]),
            "init_sigma": float(s.compute_sigma(B, J))}
    out["meta_json"] = np.array(json.dumps(meta))
    OUT.mkdir(exist_ok=True)
    np.savez_compressed(OUT / "digest" / f"soft64_{dtype_name}.npz", **out)
    print("digest", dtype_name, meta["init_sigma"])


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
        (OUT / "digest").mkdir(parents=True, exist_ok=True)
        for dtype_name in ("float64", "float32"):
            if only and f"soft64_{dtype_name}" not in only:
                continue
            run_digest_case(ref, torch, dtype_name)


if __name__ == "__main__":
    main()
