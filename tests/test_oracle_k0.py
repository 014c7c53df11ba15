"""The oracle's coupled horizontal-harmonic closures (coupled_k0,
constrained_spline_harmonic) against the reference's vectors
(tests/golden/k0/*.npz, written by tests/golden/make_k0.py), plus pins of the
closure algebra that do not need the reference: the profile solves satisfy
their normal equations and constraints, and the B-spline basis is a
partition of unity."""
import glob
import json
import os

import numpy as np
import pytest

from oracle.nfgs_oracle import OracleNFGS
from tests.golden_cases import GOLDEN_DIR, rel_l2

CASES = sorted(glob.glob(os.path.join(GOLDEN_DIR, "k0", "*.npz")))


def load_k0(path):
    g = dict(np.load(path))
    return g, json.loads(str(g.pop("meta_json")))


def oracle_for(meta):
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    return OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=np.dtype(c["dtype"]),
                      bottom_divergence_robin_weight=c["rho"], source_zero_mode_policy="project",
                      d0_zero_mode_policy="project", **c["closure"])


def _h_err(a, b, scale):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))) / scale


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_oracle_closure_matches_reference(path):
    g, meta = load_k0(path)
    c = meta["case"]
    tol = 1e-10 if c["dtype"] == "float64" else 2e-4
    o = oracle_for(meta)
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    scale = float(np.max(np.abs(g["in_bz"])))
    assert _h_err(o.H_target, g["H_bottom_target"], scale) < tol
    assert _h_err(o.H, g["H_k0_setup"], scale) < tol
    st = o.initial_state()
    for k, v in zip(("B", "J", "Phi_c", "Az"), st):
        assert rel_l2(v, g["init_" + k]) < tol, "init_" + k
    for step in (1, 2):
        st = o.step(*st)
        for k, v in zip(("B", "J", "Phi_c", "Az"), st):
            assert rel_l2(v, g[f"step{step}_{k}"]) < tol, f"step{step}_{k}"
        assert _h_err(o.H, g[f"step{step}_H_k0"], scale) < tol, f"step{step}_H"
        assert abs(o.sigma(st[0], st[1]) - float(g[f"step{step}_sigma"])) < tol
        ref_diag = meta[f"step{step}_k0"]
        for key in ("ode_residual_relative", "face_ode_residual_relative",
                    "desired_mean_current_rms", "profile_max_abs"):
            if key in ref_diag:
                assert o.k0_diag[key] == pytest.approx(ref_diag[key], rel=1e-6 if tol < 1e-6
                                                       else 1e-2, abs=1e-12), key
    o2 = oracle_for(meta)
    o2.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    it = meta["iterate"]
    h = o2.iterate(max_iter=it["max_iter"], tol=it["tol"], patience=it["patience"],
                   sigma_improvement_tol=it["sigma_improvement_tol"])
    assert h["stop_reason"] == meta["hist"]["stop_reason"]
    assert h["best_iter"] == meta["hist"]["best_iter"]
    for key in ("sigma", "divB", "energy"):
        assert rel_l2(h[key], g["hist_" + key]) < tol, key
    assert _h_err(o2.H, g["final_H_k0"], scale) < tol
    out = o2.fields_to_numpy()
    for k, v in out.items():
        assert rel_l2(v, g["out_" + k]) < tol, k


@pytest.mark.parametrize("policy,disc,smooth", [("coupled_k0", "staggered_face", 1e-3),
                                                ("coupled_k0", "legacy_plane_centred", 1e-2),
                                                ("constrained_spline_harmonic", None, 1e-3),
                                                ("constrained_spline_harmonic", None, 0.0)])
def test_profile_solves_satisfy_their_normal_equations(policy, disc, smooth):
    """The least-squares profile with its bottom value fixed: the residual
    of the design is orthogonal to the design's columns (normal equations),
    and the spline meets its two exact constraints."""
    kw = dict(horizontal_zero_mode_policy=policy, k0_ode_smoothness=smooth)
    if disc:
        kw["coupled_k0_discretization"] = disc
    o = OracleNFGS(12, 10, 17, 1.0, 1.0, 0.7, **kw)
    o.Nz = 17
    o._k0_setup()
    rng = np.random.default_rng(0)
    q = rng.normal(size=o.Nz)
    h0 = 0.37
    p = o._k0_profile(q, h0)
    assert p[0] == pytest.approx(h0, abs=1e-13)
    m = o.k0_mats
    if policy == "constrained_spline_harmonic":
        basis, design = m["basis"], m["design"]
        coef = np.linalg.lstsq(basis, p, rcond=None)[0]
        rhs = np.concatenate([0.5 * (q[:-1] + q[1:]), np.zeros(design.shape[0] - o.Nz + 1)])
        grad = design.T @ (design @ coef - rhs)
        cons = m["kkt"][-2:, :o.k0_nb]
        # constraint rows: exact bottom value and zero top derivative
        assert cons[0] @ coef == pytest.approx(h0, abs=1e-12)
        assert cons[1] @ coef == pytest.approx(0.0, abs=1e-10)
        # stationarity: the gradient lies in the span of the constraint rows
        lam = np.linalg.lstsq(cons.T, -grad, rcond=None)[0]
        assert np.linalg.norm(grad + cons.T @ lam) < 1e-8 * max(1.0, np.linalg.norm(grad))
    else:
        A = m["A"]
        D, C = m["D"], m["C"]
        s = np.sqrt(smooth)
        if disc == "legacy_plane_centred":
            b = [q[1:] - D[1:, 0] * h0, -s * C[1:, 0] * h0]
        else:
            b = [0.5 * (q[:-1] + q[1:]) - D[:, 0] * h0, -s * C[:, 0] * h0]
        r = A @ p[1:] - np.concatenate(b)
        assert np.linalg.norm(A.T @ r) < 1e-9 * max(1.0, np.linalg.norm(A.T @ np.concatenate(b)))


def test_bspline_basis_is_a_partition_of_unity():
    for nb in (4, 6, 8, 12):
        b = OracleNFGS.bspline_basis(33, nb)
        np.testing.assert_allclose(b.sum(axis=1), 1.0, atol=1e-14)
        assert (b >= -1e-15).all()
        assert b[0, 0] == 1.0 and b[-1, -1] == 1.0
