"""The oracle's soft_cfcf start against the reference's vectors
(tests/golden/soft/*.npz, written by tests/golden/make_soft.py), plus pins
that do not need the reference: the essential rows hold exactly, the
profile minimises its modal form, and at eps = 0 the Az profile is the
closed-form solution of L_k a = 0 that meets the bottom value and the top
Robin row."""
import glob
import json
import os

import numpy as np
import pytest

from oracle.nfgs_oracle import OracleNFGS
from tests.golden_cases import GOLDEN_DIR, rel_l2

CASES = sorted(glob.glob(os.path.join(GOLDEN_DIR, "soft", "*.npz")))

# diagnostics that measure exact rows: their values are rounding noise
ROUNDOFF_KEYS = ("bottom_phi_rms", "bottom_dphi_rms", "bottom_az_rms", "bottom_bz_rms",
                 "bottom_jz_rms", "bottom_bh_current_carrying_rms", "top_phi_transparent_rms",
                 "top_az_transparent_rms")


def load_soft(path):
    g = dict(np.load(path))
    return g, json.loads(str(g.pop("meta_json")))


def oracle_for(meta):
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    return OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=np.dtype(c["dtype"]),
                      bottom_divergence_robin_weight=c["rho"], source_zero_mode_policy="project",
                      d0_zero_mode_policy="project", **c["extra"])


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_oracle_soft_start_matches_reference(path):
    g, meta = load_soft(path)
    fp64 = meta["case"]["dtype"] == "float64"
    tol = 5e-12 if fp64 else 2e-5
    o = oracle_for(meta)
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    assert o.soft_eps == pytest.approx(meta["soft_cfcf_epsilon"], rel=1e-15)
    B, J, phi, az = o.soft_initial_state()
    ph, ah, ak = o.soft_profiles
    assert rel_l2(ak, g["az_k"]) < tol
    assert rel_l2(ph, g["phi_hat"]) < tol
    assert rel_l2(ah, g["az_hat"]) < tol
    for k, v in zip(("B", "J", "Phi_c", "Az"), (B, J, phi, az)):
        assert rel_l2(v, g["init_" + k]) < tol, k
    assert o.sigma(B, J) == pytest.approx(float(g["init_sigma"]), rel=tol)
    d = o.soft_diagnostics(B, J, phi, az)
    for key, ref in meta["soft"].items():
        if key in ROUNDOFF_KEYS:
            assert abs(d[key]) < (1e-12 if fp64 else 1e-5) and abs(ref) < (1e-12 if fp64 else 1e-5)
        else:
            assert d[key] == pytest.approx(ref, rel=1e-10 if fp64 else 1e-5), key


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_oracle_soft_trajectory_matches_reference(path):
    g, meta = load_soft(path)
    fp64 = meta["case"]["dtype"] == "float64"
    tol = 1e-10 if fp64 else 1e-3
    o = oracle_for(meta)
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    it = meta["iterate"]
    h = o.iterate(max_iter=it["max_iter"], tol=it["tol"], patience=it["patience"],
                  sigma_improvement_tol=it["sigma_improvement_tol"], start="soft_cfcf")
    assert h["stop_reason"] == meta["hist"]["stop_reason"]
    if fp64:
        assert h["best_iter"] == meta["hist"]["best_iter"]
    for key in ("sigma", "divB", "energy"):
        assert rel_l2(h[key], g["hist_" + key]) < tol, key
    if fp64:
        for k, v in o.fields_to_numpy().items():
            assert rel_l2(v, g["out_" + k]) < tol, k


def _modal_objective(o, p, k, kind):
    """The modal form the start minimises, evaluated directly on a profile."""
    dz, eps = o.dz, o.soft_eps
    lap = (p[2:] - 2.0 * p[1:-1] + p[:-2]) / dz ** 2 - k * k * p[1:-1]
    if kind == "phi":
        d = np.empty_like(lap)
        d[0] = (-3.0 * lap[0] + 4.0 * lap[1] - lap[2]) / (2.0 * dz)
        d[1:-1] = (lap[2:] - lap[:-2]) / (2.0 * dz)
        d[-1] = (3.0 * lap[-1] - 4.0 * lap[-2] + lap[-3]) / (2.0 * dz)
        return float((d ** 2).sum() + (k * k + eps) * (lap ** 2).sum())
    dp = (p[2:] - p[:-2]) / (2.0 * dz)
    return float((lap ** 2).sum() + eps * ((dp ** 2).sum() + k * k * (p[1:-1] ** 2).sum()))


@pytest.mark.parametrize("kind", ["phi", "az"])
def test_essential_rows_hold_and_the_profile_minimises_its_form(kind):
    o = OracleNFGS(12, 10, 17, 1.0, 1.0, 0.7)
    o.Nz = 17
    rng = np.random.default_rng(3)
    k = np.array([0.05, 0.3, 1.1, 2.7])
    b = rng.normal(size=k.size)
    P = o.soft_modes(k, b, kind)
    dz = o.dz
    np.testing.assert_allclose(P[0], b, rtol=0, atol=1e-15)
    top = (3.0 * P[-1] - 4.0 * P[-2] + P[-3]) / (2.0 * dz) + k * P[-1]
    assert np.abs(top).max() < 1e-11 * max(1.0, np.abs(P).max())
    if kind == "phi":
        bottom = (-3.0 * P[0] + 4.0 * P[1] - P[2]) / (2.0 * dz) - k * P[0]
        assert np.abs(bottom).max() < 1e-11 * max(1.0, np.abs(P).max())
    # any admissible change of the free values raises the form: move the
    # free planes and re-close the dependent ones exactly as the rows demand
    first = 2 if kind == "phi" else 1
    for m in range(k.size):
        f0 = _modal_objective(o, P[:, m], k[m], kind)
        for _ in range(5):
            q = P[:, m].copy()
            q[first:-1] += 1e-4 * rng.normal(size=q[first:-1].size)
            if kind == "phi":
                q[1] = 0.25 * q[2] + (0.75 + 0.5 * dz * k[m]) * q[0]
            den = 3.0 + 2.0 * dz * k[m]
            q[-1] = (4.0 * q[-2] - q[-3]) / den
            assert _modal_objective(o, q, k[m], kind) >= f0 * (1.0 - 1e-12)


def test_az_profile_at_eps0_is_the_closed_form_harmonic_mode():
    """With eps = 0 the Az rows are L_k a = 0 on the interior planes; the two
    roots r, 1/r of r + 1/r = 2 + k^2 dz^2 span its solutions, and the
    bottom value with the top Robin row fixes the combination.  That profile
    makes every row zero, so it is the minimiser."""
    o = OracleNFGS(12, 10, 13, 1.0, 1.0, 0.5, soft_cfcf_epsilon=0.0)
    o.Nz = 13
    N, dz = o.Nz, o.dz
    k = np.array([0.2, 0.9, 2.5])
    b = np.array([1.0, -0.4, 2.0])
    P = o.soft_modes(k, b, "az")
    z = np.arange(N)
    for m in range(k.size):
        c = 2.0 + (k[m] * dz) ** 2
        r = (c - np.sqrt(c * c - 4.0)) / 2.0
        basis = np.stack([r ** z, r ** (-z)], axis=1)
        rob = lambda v: (3.0 * v[-1] - 4.0 * v[-2] + v[-3]) / (2.0 * dz) + k[m] * v[-1]
        M = np.array([[basis[0, 0], basis[0, 1]], [rob(basis[:, 0]), rob(basis[:, 1])]])
        coef = np.linalg.solve(M, np.array([b[m], 0.0]))
        np.testing.assert_allclose(P[:, m], basis @ coef, rtol=1e-9, atol=1e-12)


def test_modes_without_wavenumber_are_zero():
    o = OracleNFGS(12, 10, 9, 1.0, 1.0, 1.0)
    o.Nz = 9
    P = o.soft_modes(np.array([0.0, 0.5]), np.array([3.0, 1.0]), "phi")
    assert np.all(P[:, 0] == 0.0) and np.any(P[:, 1] != 0.0)
