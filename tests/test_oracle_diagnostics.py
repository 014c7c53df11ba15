"""The oracle's per-iteration diagnostics and fixed-point stop against the
reference's golden vectors (tests/golden/diag/*.npz, written by
tests/golden/make_diagnostics.py with record_iteration_diagnostics=True).

Compared: the diagnostics dict of two iteration_step calls from the gs_cfcf
state, every diagnostic list of a 30-step iterate() (sigma_best), and a
fixed_point run (its stop, the criteria flags per step, the streak and the
selected final iterate).  Diagnostics that are genuine quantities (map
residuals, the B_c,z and J_z source residuals, the bottom imprint, the face
fractions) match at 1e-10 in FP64 (FP32: 5e-3, the source residuals 5e-2).
The scalar PDE residuals of an exact
solve are rounding noise in both codes (~1e-15 in FP64): they are held below
1e-10 (FP64) and 1e-4 (FP32) instead of being compared.
"""
import os

import numpy as np
import pytest

from oracle.nfgs_oracle import OracleNFGS
from tests.diag_cases import CASES, check_diag, load


def oracle_for(meta):
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    return OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=np.dtype(c["dtype"]),
                      bottom_divergence_robin_weight=c["rho"], source_zero_mode_policy="project",
                      d0_zero_mode_policy="project", **dict(c["extra"]))


def test_fixtures_present():
    assert len(CASES) == 4


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_step_diagnostics_match_the_reference(path):
    g, meta = load(path)
    fp64 = meta["case"]["dtype"] == "float64"
    o = oracle_for(meta)
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    st = o.initial_state()
    for i, want in enumerate(meta["steps"], start=1):
        B, J, phi, az, got = o.step_diagnostics(*st)
        st = (B, J, phi, az)
        check_diag(got, want, fp64, f"step {i}")
        if fp64:
            assert np.max(np.abs(phi - g[f"step{i}_Phi_c"])) < 1e-10 * np.max(np.abs(phi))


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_iterate_with_diagnostics_matches_the_reference(path):
    g, meta = load(path)
    c = meta["case"]
    fp64 = c["dtype"] == "float64"
    it = c["iterate"]
    o = oracle_for(meta)
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    h = o.iterate(max_iter=it["max_iter"], tol=it["tol"], patience=it["patience"],
                  sigma_improvement_tol=it["sigma_improvement_tol"], record_diagnostics=True,
                  stop_mode=it["stop_mode"])
    assert h["stop_reason"] == meta["stop_reason"]
    assert h["final_iter"] == meta["final_iter"] and h["best_iter"] == meta["best_iter"]
    assert h["selected_state"] == meta["selected_state"]
    n = len(g["hist_sigma"])
    assert len(h["sigma"]) == n
    for k in OracleNFGS.DIAGNOSTIC_KEYS:
        assert len(h[k]) == n
        for i in range(n):
            check_diag({k: h[k][i]}, {k: g["hist_" + k][i]}, fp64, f"iteration {i + 1}",
                       drift=10.0)
    if it["stop_mode"] == "fixed_point":
        assert h["fixed_point_converged"] == meta["fixed_point_converged"] is True
        assert h["fixed_point_streak"] == meta["fixed_point_streak"]
        assert list(h["fixed_point_criteria_met"]) == [int(v) for v in g["hist_fixed_point_criteria_met"]]
    f = o.fields_to_numpy()
    tol = 1e-10 if fp64 else 5e-4
    for k in ("Bx", "By", "Bz"):
        a, b = f[k].astype(np.float64), g["out_" + k].astype(np.float64)
        assert np.linalg.norm(a - b) <= tol * np.linalg.norm(b), k
