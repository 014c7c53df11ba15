"""The coupled horizontal-harmonic closures on the GPU (coupled_k0 with both
discretisations, constrained_spline_harmonic) against the reference's vectors
in tests/golden/k0/ (written by tests/golden/make_k0.py): the setup profile,
the gs_cfcf state, two iteration_step() states with the profile after each,
a 30-step iterate() trajectory, the selected state's profile and fields, and
the closure diagnostics."""
import glob
import json
import os
import sys

import numpy as np
import pytest
import torch

from tests.golden_cases import GOLDEN_DIR, rel_l2

pytestmark = pytest.mark.gpu

CASES = sorted(glob.glob(os.path.join(GOLDEN_DIR, "k0", "*.npz")))


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")


def _load(path):
    g = dict(np.load(path))
    return g, json.loads(str(g.pop("meta_json")))


def _solver(meta, **extra):
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
    kw.update(c["closure"])
    kw.update(extra)
    return NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, **kw)


def _set(s, g):
    s.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"], zero_net_current=False,
                                 smooth_passes=0, ijzbcor=3)


def _t(x):
    return x.detach().double().cpu().numpy()


def _h_err(a, b, scale):
    return float(np.max(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)))) / scale


# FP64: the 1e-10 relative bar.  FP32 against the reference's own FP32 run:
# 2e-4 for the states of the first steps (the plain solver's FP32 bar for
# currents and scalars), 2e-3 after 30 coupled steps, where the profile's
# least-squares solve feeds roundoff back into every later step.
def _tol(meta, late=False):
    if meta["case"]["dtype"] == "float64":
        return 1e-10
    return 2e-3 if late else 2e-4


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_closure_steps_match_reference(path):
    g, meta = _load(path)
    _announce(f"k0 steps {os.path.basename(path)}")
    tol = _tol(meta)
    s = _solver(meta)
    _set(s, g)
    assert tuple(g["calc_shape_zyx"]) == (s.Nz, s.Ny, s.Nx)
    scale = float(np.max(np.abs(g["in_bz"])))
    assert _h_err(_t(s._H_k0_bottom_target), g["H_bottom_target"], scale) < tol
    assert _h_err(_t(s._H_k0), g["H_k0_setup"], scale) < tol
    st = s.compute_initial_conditions(init_mode="gs_cfcf")
    for k, v in zip(("B", "J", "Phi_c", "Az"), st[:4]):
        assert rel_l2(_t(v), g["init_" + k]) < tol, "init_" + k
    assert _h_err(_t(s._H_k0), g["init_H_k0"], scale) < tol
    for step in (1, 2):
        st = s.iteration_step(*st[:4], mu=1.0)
        for k, v in zip(("B", "J", "Phi_c", "Az"), st[:4]):
            assert rel_l2(_t(v), g[f"step{step}_{k}"]) < tol, f"step{step}_{k}"
        assert _h_err(_t(s._H_k0), g[f"step{step}_H_k0"], scale) < tol, f"step{step}_H"
        sigma = s.compute_sigma(st[0], st[1])
        assert abs(sigma - float(g[f"step{step}_sigma"])) < tol * max(1.0, abs(sigma))
        ref = meta[f"step{step}_k0"]
        for key in ("ode_residual_relative", "face_ode_residual_relative",
                    "desired_mean_current_rms", "profile_max_abs", "gridscale_profile_relative"):
            if key in ref:
                assert s._k0_diagnostics[key] == pytest.approx(
                    ref[key], rel=1e-6 if tol < 1e-6 else 1e-2, abs=1e-12), key


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_closure_iterate_matches_reference(path):
    g, meta = _load(path)
    _announce(f"k0 iterate {os.path.basename(path)}")
    tol = _tol(meta, late=True)
    s = _solver(meta)
    _set(s, g)
    it = dict(meta["iterate"])
    h = s.iterate(**it)
    ref = meta["hist"]
    assert h["stop_reason"] == ref["stop_reason"]
    if meta["case"]["dtype"] == "float64":
        assert h["best_iter"] == ref["best_iter"]
    for key in ("sigma", "divB", "energy"):
        assert rel_l2(h[key], g["hist_" + key]) < tol, key
    scale = float(np.max(np.abs(g["in_bz"])))
    assert _h_err(_t(s._H_k0), g["final_H_k0"], scale) < tol
    out = s.fields_to_numpy()
    for k, v in out.items():
        assert rel_l2(v, g["out_" + k]) < tol, k
    # the last step's closure diagnostics and the running maximum
    fin = meta["final_k0"]
    kd = h["init_diagnostics"]["coupled_k0"]
    for key in ("ode_residual_relative", "ode_residual_relative_max",
                "face_ode_residual_relative", "desired_mean_current_rms"):
        assert kd[key] == pytest.approx(fin[key], rel=1e-6 if tol < 1e-6 else 2e-2), key
    # source diagnostics carry the selected profile under the policy's name
    policy = meta["case"]["closure"]["horizontal_zero_mode_policy"]
    prof = np.array(h["source_diagnostics"][policy]["profile_by_height"])
    assert _h_err(prof, g["final_H_k0"], scale) < tol


def test_spline_current_on_the_walls_is_the_harmonic_current():
    """Under the spline closure the side walls carry (-dH_y/dz, dH_x/dz) of the
    state's profile and J_z is the face-averaged PT current; under coupled_k0
    the walls carry no current."""
    path = os.path.join(GOLDEN_DIR, "k0", "spline_rho05.npz")
    g, meta = _load(path)
    s = _solver(meta)
    _set(s, g)
    s.iterate(**meta["iterate"])
    J = _t(s._J)
    H = _t(s._H_k0)
    dz = s.dz
    d = np.zeros_like(H)
    face = (H[:, 1:] - H[:, :-1]) / dz
    d[:, 1:-1] = 0.5 * (face[:, :-1] + face[:, 1:])
    d[:, 0] = (-3.0 * H[:, 0] + 4.0 * H[:, 1] - H[:, 2]) / (2.0 * dz)
    for k in range(s.Nz - 1):
        for wall in (J[:, k, :, 0], J[:, k, :, -1], J[:, k, 0, :], J[:, k, -1, :]):
            np.testing.assert_allclose(wall[0], -d[1, k], rtol=1e-12, atol=1e-14)
            np.testing.assert_allclose(wall[1], d[0, k], rtol=1e-12, atol=1e-14)
    assert np.all(J[:, -1] == 0.0)
    gc = os.path.join(GOLDEN_DIR, "k0", "coupled_rho05.npz")
    g2, meta2 = _load(gc)
    s2 = _solver(meta2)
    _set(s2, g2)
    s2.iterate(**meta2["iterate"])
    J2 = _t(s2._J)
    assert np.all(J2[:2, :, :, 0] == 0.0) and np.all(J2[:2, :, 0, :] == 0.0)


def test_shared_boundary_starts_from_the_setup_profile():
    """share_boundary() under a closure gives the second solver the setup
    profile and its own records: its solve equals a fresh solver's."""
    path = os.path.join(GOLDEN_DIR, "k0", "coupled_rho05.npz")
    g, meta = _load(path)
    a = _solver(meta)
    _set(a, g)
    ha = a.iterate(**meta["iterate"])
    b = _solver(meta)
    b.share_boundary(a)
    hb = b.iterate(**meta["iterate"])
    assert hb["sigma"] == ha["sigma"]
    assert torch.equal(a._B, b._B)


def test_closure_options_that_are_not_built_raise():
    path = os.path.join(GOLDEN_DIR, "k0", "coupled_rho05.npz")
    _, meta = _load(path)
    with pytest.raises(NotImplementedError):
        _solver(meta, record_iteration_diagnostics=True)
