"""The oracle's relaxed (mu < 1), safeguarded (adaptive mu with rollbacks)
and polarity-masked iterations against the reference's vectors
(tests/golden/relax/*.npz, written by tests/golden/make_relax.py), plus pins
that do not need the reference: mu = 1 is the plain step exactly, and a
relaxed step is the blend of the plain step's scalars with the old ones."""
import glob
import json
import os

import numpy as np
import pytest

from oracle.nfgs_oracle import OracleNFGS
from tests.golden_cases import GOLDEN_DIR, rel_l2

CASES = sorted(glob.glob(os.path.join(GOLDEN_DIR, "relax", "*.npz")))


def load_relax(path):
    g = dict(np.load(path))
    return g, json.loads(str(g.pop("meta_json")))


def oracle_for(meta):
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    return OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=np.dtype(c["dtype"]),
                      bottom_divergence_robin_weight=c["rho"], source_zero_mode_policy="project",
                      d0_zero_mode_policy="project", **c["extra"])


def iterate_kwargs(meta):
    it = dict(meta["iterate"])
    kw = {k: it[k] for k in ("max_iter", "tol", "patience", "sigma_improvement_tol", "mu")}
    for k in ("adaptive_mu", "mu_min", "mu_max", "mu_backoff", "mu_growth", "growth_streak",
              "divergence_factor", "max_rollbacks"):
        if k in it:
            kw[k] = it[k]
    kw["polarity"] = it.get("polarity_mask", "both")
    return kw


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_oracle_relaxed_steps_and_trajectory_match_reference(path):
    g, meta = load_relax(path)
    fp64 = meta["case"]["dtype"] == "float64"
    tol = 1e-10 if fp64 else 2e-4
    o = oracle_for(meta)
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    st = o.initial_state()
    for k, v in zip(("B", "J", "Phi_c", "Az"), st):
        assert rel_l2(v, g["init_" + k]) < tol, "init_" + k
    sk = meta["step"]
    for step in (1, 2):
        st = o.step(*st, mu=sk["mu"], polarity=sk.get("polarity_mask", "both"))
        for k, v in zip(("B", "J", "Phi_c", "Az"), st):
            assert rel_l2(v, g[f"step{step}_{k}"]) < tol, f"step{step}_{k}"
    o2 = oracle_for(meta)
    o2.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    h = o2.iterate(**iterate_kwargs(meta))
    ref = meta["hist"]
    assert h["stop_reason"] == ref["stop_reason"]
    assert h["rollbacks"] == ref["rollbacks"]
    assert len(h["sigma"]) == len(g["hist_sigma"])
    np.testing.assert_array_equal(np.array(h["rollback"], dtype=float), g["hist_rollback"])
    np.testing.assert_allclose(h["mu"], g["hist_mu"], rtol=1e-12)
    if fp64:
        assert h["best_iter"] == ref["best_iter"]
    tt = 1e-10 if fp64 else 1e-3
    ok = ~np.isnan(g["hist_divB"])
    for key in ("sigma", "divB", "energy"):
        a = np.array(h[key], dtype=np.float64)
        np.testing.assert_array_equal(np.isnan(a), np.isnan(g["hist_" + key]))
        assert rel_l2(a[ok], g["hist_" + key][ok]) < tt, key
    if fp64:
        for k, v in o2.fields_to_numpy().items():
            assert rel_l2(v, g["out_" + k]) < tt, k


def test_mu_one_is_the_plain_step_and_mu_blends_the_scalars():
    path = os.path.join(GOLDEN_DIR, "relax", "mu05.npz")
    g, meta = load_relax(path)
    o = oracle_for(meta)
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    st = o.initial_state()
    plain = o.step(*st)
    one = o.step(*st, mu=1.0)
    for a, b in zip(plain, one):
        assert np.array_equal(a, b)
    mu = 0.3
    relaxed = o.step(*st, mu=mu)
    for new, star, old in ((relaxed[2], plain[2], st[2]), (relaxed[3], plain[3], st[3])):
        blend = mu * star + (1.0 - mu) * old
        np.testing.assert_allclose(new[1:-1, 1:-1, 1:-1], blend[1:-1, 1:-1, 1:-1], rtol=0,
                                   atol=1e-14 * np.abs(blend).max())
