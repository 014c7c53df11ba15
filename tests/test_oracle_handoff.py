"""The oracle's caller-field start (init_mode="field").

project_field restates the cfcf handoff of nfgs_solver.py:6080-6147 with the
field and Az given by the caller: Phi_c from the skip-bottom endpoint-Neumann
Poisson of Bc_z (:1120-1130, :1146-1180), the constant mode of every plane by
height (:4466-4468), then the state's B and J.  The reference has no such
entry point, so these tests pin it by what the mathematics fixes: on a field
the NFGS scalars can represent, the projection returns that field and its
current exactly, and iterating from it continues the trajectory it came from.
"""
import os

import numpy as np
import pytest

from oracle import nfgs_oracle as O
from tests.golden_cases import case_files, load, rel_l2

CASE = [p for p in case_files() if os.path.basename(p) == "mudpack_rho05.npz"][0]


def _oracle():
    g, meta = load(CASE)
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    o = O.OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=np.float64,
                     bottom_divergence_robin_weight=c["rho"], pad_calculation_domain=c["pad"],
                     calculation_domain_mode=c["mode"], source_zero_mode_policy="project",
                     d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    return o


def test_projection_returns_a_representable_state():
    o = _oracle()
    st = o.initial_state()
    B, J, phi, az = o.step(*o.step(*st))
    phi2, az2, mean, Bp, Jp = o.project_field(B, az)
    # the field and its current come back to roundoff
    assert rel_l2(Bp, B) < 1e-12
    assert rel_l2(Jp, J) < 1e-11
    np.testing.assert_array_equal(az2, az)
    # Phi_c is fixed up to one constant per plane above the bottom
    assert rel_l2(phi2[0], phi[0]) == 0.0
    d = (phi2 - phi)[1:]
    spread = d.max(axis=(1, 2)) - d.min(axis=(1, 2))
    assert np.max(spread) < 1e-12 * np.max(np.abs(phi))
    # lap_xy of a mirror-Neumann field and B_z above the bottom have no constant mode
    assert np.max(np.abs(mean)) < 1e-12 * np.max(np.abs(B[2]))


def test_projection_drops_and_reports_the_constant_mode():
    o = _oracle()
    B, J, phi, az = o.step(*o.initial_state())
    shift = np.linspace(0.5, 2.0, o.Nz)
    Bs = B.copy()
    Bs[2] += shift[:, None, None]
    _, _, mean, Bp, _ = o.project_field(Bs, az)
    _, _, mean0, Bp0, _ = o.project_field(B, az)
    np.testing.assert_allclose(mean - mean0, shift, rtol=1e-12)
    assert rel_l2(Bp, Bp0) < 1e-13


def test_projection_imposes_the_fixed_planes_of_az():
    o = _oracle()
    B, J, phi, az = o.step(*o.initial_state())
    rng = np.random.default_rng(0)
    noisy = az + rng.standard_normal(az.shape)
    _, az2, _, _, _ = o.project_field(B, noisy)
    np.testing.assert_array_equal(az2[0], o.Az_bottom)
    assert not np.any(az2[-1])
    for a in (az2[1:, 0, :], az2[1:, -1, :], az2[1:, :, 0], az2[1:, :, -1]):
        assert not np.any(a)
    np.testing.assert_array_equal(az2[1:-1, 1:-1, 1:-1], noisy[1:-1, 1:-1, 1:-1])


def test_iterate_from_a_state_continues_its_trajectory():
    o = _oracle()
    h_ref = o.iterate(max_iter=6, patience=50, sigma_improvement_tol=None, tol=0.0)
    o2 = _oracle()
    st = o2.initial_state()
    for _ in range(2):
        st = o2.step(*st)
    B, J, phi, az = st
    h = o2.iterate(max_iter=4, patience=50, sigma_improvement_tol=None, tol=0.0, start=(B, az))
    np.testing.assert_allclose(h["sigma"], h_ref["sigma"][2:6], rtol=1e-10)
    assert h["initial_sigma"] == pytest.approx(h_ref["sigma"][1], rel=1e-10)
