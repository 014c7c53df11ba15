"""The oracle's two-stage poloidal update (poloidal_update_mode="two_stage").

solve_phi_c_two_stage restates nfgs_solver.py:5674-5725: the modal Robin solve
for B_c,z with the undivided source and the bottom data scaled by lambda, then
Phi_c_hat = DCT2(B_c,z) / lambda per plane.  Since (d_zz + lambda) is linear
and the Robin rows are homogeneous in the unknown, B_c,z = lambda Phi_c mode
by mode: the two-stage update equals the direct one (:5627-5672) up to
rounding.  The reference's tests hold no two-stage vector; these pins are that
identity (FP64, a whole short iteration) and the bottom plane being the
prescribed Phi_above_bottom exactly.
"""
import os

import numpy as np
import pytest

from oracle import nfgs_oracle as O
from tests.golden_cases import case_files, load, rel_l2


def _pair(name, mode):
    path = [p for p in case_files() if os.path.basename(p) == name + ".npz"][0]
    g, meta = load(path)
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    o = O.OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=np.float64,
                     bottom_divergence_robin_weight=c["rho"], pad_calculation_domain=c["pad"],
                     calculation_domain_mode=c["mode"], source_zero_mode_policy="project",
                     d0_zero_mode_policy="project", poloidal_update_mode=mode)
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    return o


@pytest.mark.parametrize("name", ["mudpack_rho05", "mudpack_rho0", "mudpack_rho1"])
def test_two_stage_equals_direct_in_fp64(name):
    d, t = _pair(name, "direct"), _pair(name, "two_stage")
    hd = d.iterate(max_iter=3, tol=1e-30, patience=50, sigma_improvement_tol=None)
    ht = t.iterate(max_iter=3, tol=1e-30, patience=50, sigma_improvement_tol=None)
    assert rel_l2(np.array(ht["sigma"]), np.array(hd["sigma"])) < 1e-11
    fd, ft = d.fields_to_numpy(), t.fields_to_numpy()
    for k in fd:
        assert rel_l2(ft[k], fd[k]) < 1e-10, k


def test_two_stage_keeps_the_bottom_plane_and_rejects_unknown_modes():
    t = _pair("mudpack_rho05", "two_stage")
    rng = np.random.default_rng(0)
    src = rng.standard_normal((t.Nz, t.Ny, t.Nx))
    phi = t.solve_phi_c(src)
    np.testing.assert_array_equal(phi[0], t.Phi_above_bottom)
    d = _pair("mudpack_rho05", "direct")
    assert rel_l2(phi, d.solve_phi_c(src)) < 1e-11
    with pytest.raises(ValueError):
        O.OracleNFGS(12, 10, 9, poloidal_update_mode="split")
