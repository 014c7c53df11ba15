"""Caller-field start (init_mode="field") on the GPU against the oracle.

nfgs_project_field plus the state pass (nfgs_amd/solver.py:_init_from_field)
against oracle.project_field (the cfcf handoff of nfgs_solver.py:6080-6147
with the field and Az given; pinned in tests/test_oracle_handoff.py), on the
reference's golden boundaries, and a warm-started iterate against the
oracle's iterate from the same state.
"""
import os
import sys

import numpy as np
import pytest
import torch

from oracle.nfgs_oracle import OracleNFGS
from tests.golden_cases import case_files, load, rel_l2

pytestmark = pytest.mark.gpu

CASES = {os.path.basename(p)[:-4]: p for p in case_files()}


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _pair(name):
    from nfgs_amd import NFGSSolverGPU
    g, meta = load(CASES[name])
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    dt = np.dtype(c["dtype"])
    o = OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=dt, bottom_divergence_robin_weight=c["rho"],
                   pad_calculation_domain=c["pad"], calculation_domain_mode=c["mode"],
                   source_zero_mode_policy="project", d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])

    def gpu(**kw):
        s = NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, device="cuda", **kw,
                          dtype=torch.float64 if dt == np.float64 else torch.float32,
                          lateral_wall="paper_bc1", buffer_thickness=0, padded_buffer_zone=True,
                          pad_calculation_domain=c["pad"], calculation_domain_mode=c["mode"],
                          bottom_divergence_robin_weight=c["rho"], mean_bz_policy="uniform_open",
                          horizontal_zero_mode_policy="frozen_linear_harmonic",
                          d0_zero_mode_policy="project", source_zero_mode_policy="project",
                          harmonic_residual_policy="diagnose", poloidal_update_mode="direct",
                          vertical_current_operator="staggered_face")
        s.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"], zero_net_current=False,
                                     smooth_passes=0, ijzbcor=3)
        return s
    return o, gpu, dt


def _np(t):
    return t.detach().double().cpu().numpy()


@pytest.mark.parametrize("name", ["mudpack_rho05", "cube16_rho05", "fp32_rho05"])
def test_projection_of_an_arbitrary_field_matches_oracle(name):
    _announce(f"field projection {name}")
    o, gpu, dt = _pair(name)
    s = gpu()
    rng = np.random.default_rng(7)
    shape = (3, o.Nz, o.Ny, o.Nx)
    B = (o.B_background + 0.3 * rng.standard_normal(shape)).astype(dt)
    Az = rng.standard_normal(shape[1:]).astype(dt)
    J = rng.standard_normal(shape).astype(dt)
    Bg, Jg, phi, az, _ = s.compute_initial_conditions(init_mode="field", init_field=B, init_az=Az,
                                                      init_current=J)
    phi_o, az_o, mean_o, Bp, Jp = o.project_field(B, Az)
    tol = dict(phi=1e-11, B=1e-11, J=1e-10) if dt == np.float64 else \
        dict(phi=2e-5, B=2e-5, J=2e-4)
    assert rel_l2(_np(phi), phi_o) < tol["phi"]
    assert rel_l2(_np(az), az_o) < tol["phi"]
    assert rel_l2(_np(Bg), Bp) < tol["B"]
    assert rel_l2(_np(Jg), Jp) < tol["J"]
    d = s._init_diagnostics
    assert d["init_mode"] == "field"
    np.testing.assert_allclose(d["mean_bcz_by_height"], mean_o, rtol=tol["phi"] * 10,
                               atol=tol["phi"] * np.max(np.abs(mean_o)))
    want_f = np.sqrt(np.mean((Bp - B) ** 2)) / np.sqrt(np.mean(B.astype(np.float64) ** 2))
    want_j = np.sqrt(np.mean((Jp - J) ** 2)) / np.sqrt(np.mean(J.astype(np.float64) ** 2))
    assert d["field_relative_difference"] == pytest.approx(want_f, rel=tol["B"] * 10)
    assert d["current_relative_difference"] == pytest.approx(want_j, rel=tol["J"] * 10)


@pytest.mark.parametrize("name", ["mudpack_rho05", "fp32_rho05"])
def test_warm_started_iterate_matches_oracle(name):
    _announce(f"warm-started iterate {name}")
    o, gpu, dt = _pair(name)
    first = gpu()
    first.iterate(max_iter=3, tol=1e-30, patience=50, sigma_improvement_tol=None, verbose=False)
    B, Az, J = first.warm_start_state()
    second = gpu()
    h = second.iterate(init_mode="field", init_field=B, init_az=Az, init_current=J, max_iter=4,
                       tol=1e-30, patience=50, sigma_improvement_tol=None, verbose=False)
    # a state of the same boundary is representable: the projection returns it
    fd = h["init_diagnostics"]["field_relative_difference"]
    assert fd < (1e-12 if dt == np.float64 else 2e-6), fd
    ho = o.iterate(max_iter=4, tol=1e-30, patience=50, sigma_improvement_tol=None,
                   start=(_np(B).astype(dt), _np(Az).astype(dt)))
    tol = 1e-10 if dt == np.float64 else 2e-4
    assert rel_l2(np.array(h["sigma"]), np.array(ho["sigma"])) < tol
    assert rel_l2(np.array(h["energy"]), np.array(ho["energy"])) < tol


def test_field_start_rejects_bad_inputs_and_applies_the_policy():
    _announce("field start errors")
    o, gpu, dt = _pair("mudpack_rho05")
    s = gpu()
    with pytest.raises(ValueError):
        s.compute_initial_conditions(init_mode="field", init_field=np.zeros((3, 2, 2, 2)),
                                     init_az=np.zeros((o.Nz, o.Ny, o.Nx)))
    with pytest.raises(ValueError):
        s.compute_initial_conditions(init_mode="field", init_field=None, init_az=None)
    with pytest.raises(RuntimeError):
        s.warm_start_state()
    # a field outside the representable subspace under handoff_projection_policy="error"
    strict = gpu(handoff_projection_policy="error", handoff_projection_tolerance=1e-6)
    rng = np.random.default_rng(3)
    B = o.B_background + rng.standard_normal(o.B_background.shape)
    with pytest.raises(ValueError, match="representable"):
        strict.compute_initial_conditions(init_mode="field", init_field=B,
                                          init_az=np.zeros((o.Nz, o.Ny, o.Nx)))
