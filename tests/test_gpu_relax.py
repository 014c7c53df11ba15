"""Relaxed (mu < 1), safeguarded (adaptive mu with rollbacks) and
polarity-masked iterations on the GPU against the reference's vectors in
tests/golden/relax/ (written by tests/golden/make_relax.py): two
iteration_step() states at the case's mu and mask, the iterate() history
(sigma, divB, energy, mu, rollback flags, stop reason) and the cropped
fields."""
import glob
import json
import os
import sys

import numpy as np
import pytest
import torch

from tests.golden_cases import GOLDEN_DIR, rel_l2

pytestmark = pytest.mark.gpu

CASES = sorted(glob.glob(os.path.join(GOLDEN_DIR, "relax", "*.npz")))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")


def _load(path):
    g = dict(np.load(path))
    return g, json.loads(str(g.pop("meta_json")))


def _solver(g, meta):
    from nfgs_amd import NFGSSolverGPU
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    dtype = torch.float64 if c["dtype"] == "float64" else torch.float32
    kw = dict(device="cuda", dtype=dtype, lateral_wall="paper_bc1", pin_potential_boundaries=False,
              buffer_thickness=0, padded_buffer_zone=True, pad_calculation_domain=True,
              calculation_domain_mode="mudpack", bottom_divergence_robin_weight=c["rho"],
              mean_bz_policy="uniform_open", d0_zero_mode_policy="project",
              source_zero_mode_policy="project", harmonic_residual_policy="diagnose",
              poloidal_update_mode="direct", vertical_current_operator="staggered_face")
    kw.update(c["extra"])
    s = NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)
    s.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"], zero_net_current=False,
                                 smooth_passes=0, ijzbcor=3)
    return s


def _t(x):
    return x.detach().double().cpu().numpy()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_relaxed_steps_match_reference(path):
    g, meta = _load(path)
    _announce(f"relax steps {os.path.basename(path)}")
    tol = 1e-10 if meta["case"]["dtype"] == "float64" else 2e-4
    s = _solver(g, meta)
    st = s.compute_initial_conditions(init_mode="gs_cfcf")
    for step in (1, 2):
        st = s.iteration_step(*st[:4], **meta["step"])
        for k, v in zip(("B", "J", "Phi_c", "Az"), st[:4]):
            assert rel_l2(_t(v), g[f"step{step}_{k}"]) < tol, f"step{step}_{k}"
        sigma = s.compute_sigma(st[0], st[1])
        assert sigma == pytest.approx(float(g[f"step{step}_sigma"]), rel=tol)


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_relaxed_iterate_matches_reference(path):
    g, meta = _load(path)
    _announce(f"relax iterate {os.path.basename(path)}")
    fp64 = meta["case"]["dtype"] == "float64"
    tol = 1e-10 if fp64 else 1e-3
    s = _solver(g, meta)
    h = s.iterate(**meta["iterate"])
    ref = meta["hist"]
    assert h["stop_reason"] == ref["stop_reason"]
    assert h["rollbacks"] == ref["rollbacks"]
    assert len(h["sigma"]) == len(g["hist_sigma"])
    np.testing.assert_array_equal(np.array(h["rollback"], dtype=float), g["hist_rollback"])
    np.testing.assert_allclose(h["mu"], g["hist_mu"], rtol=1e-12)
    if fp64:
        assert h["best_iter"] == ref["best_iter"]
    ok = ~np.isnan(g["hist_divB"])
    for key in ("sigma", "divB", "energy"):
        a = np.array(h[key], dtype=np.float64)
        np.testing.assert_array_equal(np.isnan(a), np.isnan(g["hist_" + key]))
        assert rel_l2(a[ok], g["hist_" + key][ok]) < tol, key
    for k, v in s.fields_to_numpy().items():
        assert rel_l2(v, g["out_" + k]) < (tol if fp64 else 2e-2), k


def test_mu_one_and_both_polarities_take_the_plain_step():
    """iterate(mu=1, polarity_mask='both') is the pipelined plain driver;
    adaptive_mu=True without any rollback only changes the driver, so its
    history equals the plain one while no step diverges."""
    path = os.path.join(GOLDEN_DIR, "relax", "mu05.npz")
    g, meta = _load(path)
    it = dict(meta["iterate"], mu=1.0, max_iter=20)
    a = _solver(g, meta).iterate(**it)
    b = _solver(g, meta).iterate(**dict(it, adaptive_mu=True, divergence_factor=1e9,
                                        growth_streak=10 ** 6))
    assert a["sigma"] == b["sigma"] and a["energy"] == b["energy"]
    assert b["mu"] == [1.0] * len(b["mu"]) and b["rollbacks"] == 0


def test_relaxation_arguments_are_checked():
    path = os.path.join(GOLDEN_DIR, "relax", "mu05.npz")
    g, meta = _load(path)
    s = _solver(g, meta)
    with pytest.raises(ValueError):
        s.iterate(mu=1.5, max_iter=2, verbose=False)
    with pytest.raises(ValueError):
        s.iterate(polarity_mask="up", max_iter=2, verbose=False)
    st = s.compute_initial_conditions()
    with pytest.raises(ValueError):
        s.iteration_step(*st[:4], mu=0.0)
