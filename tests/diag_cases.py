"""Shared loader and comparison for the per-iteration diagnostics fixtures
(tests/golden/diag/*.npz, written by tests/golden/make_diagnostics.py)."""
import glob
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = sorted(glob.glob(os.path.join(HERE, "golden", "diag", "*.npz")))
# the scalar PDE residuals of an exact solve are rounding noise in both codes
ROUNDOFF = ("poloidal_pde_residual_relative", "toroidal_pde_residual_relative",
            "mixed_row_residual_relative")
# constructor settings of the fixtures beside the case's own
BASE_KW = dict(lateral_wall="paper_bc1", pin_potential_boundaries=False, buffer_thickness=0,
               padded_buffer_zone=True, pad_calculation_domain=True,
               calculation_domain_mode="mudpack", mean_bz_policy="uniform_open",
               horizontal_zero_mode_policy="frozen_linear_harmonic",
               d0_zero_mode_policy="project", source_zero_mode_policy="project",
               harmonic_residual_policy="diagnose", poloidal_update_mode="direct",
               vertical_current_operator="staggered_face")


def load(path):
    g = dict(np.load(path))
    return g, json.loads(str(g.pop("meta_json")))


def check_diag(got, want, fp64, where, drift=1.0, fp64_tol=1e-10):
    """Compare diagnostics dicts.  Genuine quantities at 1e-10 relative in
    FP64 (FP32: 5e-3, times drift over a trajectory), with a floor for map
    residuals at the rounding of the state; the B_c,z and J_z source
    residuals at 1e-8 in FP64 (5e-2 in FP32): they apply fourth-order
    operators to Phi_c and Az and cancel to ~1e-3 of their source, so their
    rounding grows with the grid (1.4e-10 measured on 98x98x129); face
    fractions exactly in FP64 (FP32: 0.2% of the faces); the scalar PDE
    residuals below 1e-10 (FP64; ~1e-15 on the fixtures' grids, 2e-12 on
    98x98x129) / 1e-4 (FP32) in both.  fp64_tol raises the
    FP64 bar where the states themselves differ by more than rounding."""
    for k, v in want.items():
        g = got[k]
        if k in ROUNDOFF:
            bar = 1e-10 if fp64 else 1e-4
            assert g < bar and v < bar, (where, k, g, v)
        elif k.endswith("fraction"):
            assert abs(g - v) <= (0.0 if fp64 else 2.0 / 1000), (where, k, g, v)
        else:
            pde = "pde" in k
            tol = max(1e-8 if pde else 1e-10, fp64_tol) if fp64 else \
                (5e-2 if pde else 5e-3) * drift
            floor = 1e-13 if fp64 else 1e-6
            assert abs(g - v) <= tol * abs(v) + floor, (where, k, g, v)
