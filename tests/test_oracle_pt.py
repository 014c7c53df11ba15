"""Pins of the oracle's collocated PT reconstruction (reconstruct_pt) and its
second z derivative, which back the drop-in's per-stage methods.

reconstruct_pt restates nfgs_solver.py:3311-3341.  Its B is pinned by the
identity with the golden-fixture-pinned NFGS reconstruction (B_pt plus the
constant harmonic lift equals reconstruct()'s B), its J by the complete-PT
manufactured solution of the reference's own harness
(extended_manufactured/run.py:336-446): B and J converge to the analytic
fields at second order.
"""
import numpy as np
import pytest

from oracle import nfgs_oracle as O
from tests.golden_cases import case_files, load
from tests.manufactured import complete_fields, observed_order


def test_d2_dz2_exact_on_cubics_inside_and_at_the_bottom():
    z = np.arange(9) * 0.5
    f = (1.0 + 2.0 * z - 0.5 * z ** 2 + 0.25 * z ** 3)[:, None, None] * np.ones((1, 3, 4))
    d2 = O.d2_dz2(f, 0.5)
    want = (-1.0 + 1.5 * z)[:, None, None]
    np.testing.assert_allclose(d2[:-1], np.broadcast_to(want[:-1], d2[:-1].shape), atol=1e-12)
    # the top row is the mirror form (-2 f[-1] + 2 f[-2]) / dz^2
    np.testing.assert_allclose(d2[-1], (-2 * f[-1] + 2 * f[-2]) / 0.25, atol=1e-12)


def test_pt_field_plus_lift_is_the_nfgs_field():
    path = [p for p in case_files() if p.endswith("mudpack_rho05.npz")][0]
    g, meta = load(path)
    c = meta["case"]
    o = O.OracleNFGS(*c["observed"], 1.0, 1.0, 1.0, dtype=np.float64,
                     bottom_divergence_robin_weight=c["rho"], source_zero_mode_policy="project",
                     d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    B, J, phi, az = o.initial_state()
    st = o.step(B, J, phi, az)
    phi, az = st[2], st[3]
    bcz = O.lap_xy(phi, o.dx, o.dy)
    Bp, Jp, Bc = O.reconstruct_pt(phi, az, bcz, o.B_background, o.Bz_obs, o.Jz_T, o.dx, o.dy,
                                  o.dz)
    Bn, _ = o.reconstruct(phi, az, bcz)
    lift = Bp.copy()
    lift[0] += o.H[0][:, None, None]
    lift[1] += o.H[1][:, None, None]
    lift[2, 0] = o.Bz_obs
    np.testing.assert_allclose(lift, Bn, rtol=0, atol=1e-12 * np.abs(Bn).max())
    np.testing.assert_array_equal(Bc[2], bcz)
    np.testing.assert_array_equal(Jp[2, 0], o.Jz_T)
    assert np.all(Jp[:, -1] == 0) and np.all(Jp[:, 1:, :, 0] == 0) and np.all(Jp[:, 1:, 0] == 0)


def test_pt_reconstruction_converges_to_the_analytic_state():
    eb, ej, hs = [], [], []
    for n in (17, 33, 65):
        h = 1.0 / (n - 1)
        f = complete_fields(n, n, n, h, h, h)
        phi, az = f["phi"], f["az"]
        bcz = O.lap_xy(phi, h, h)
        B, J, _ = O.reconstruct_pt(phi, az, bcz, np.zeros((3, n, n, n)), bcz[0],
                                   -O.lap_xy(az, h, h)[0], h, h, h)
        core = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
        eb.append(np.sqrt(np.mean((B - f["b_ref"])[core] ** 2)))
        ej.append(np.sqrt(np.mean((J - f["j_ref"])[core] ** 2)))
        hs.append(h)
    assert min(observed_order(eb, hs)) > 1.8, eb
    assert min(observed_order(ej, hs)) > 1.8, ej
    # a planted sign error in the J reconstruction (d/dy L with the wrong
    # sign) is caught: the error then does not converge
    n = 33
    h = 1.0 / (n - 1)
    f = complete_fields(n, n, n, h, h, h)
    bcz = O.lap_xy(f["phi"], h, h)
    lphi = bcz + O.d2_dz2(f["phi"], h)
    wrong = O.d_dx(O.d_dz(f["az"], h), h) - O.d_dy(lphi, h)
    core = (slice(1, -1), slice(1, -1), slice(1, -1))
    assert np.sqrt(np.mean((wrong - f["j_ref"][0])[core] ** 2)) > 10 * ej[1]
