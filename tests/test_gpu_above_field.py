"""The outward Gaussian field B^> on z = 0 on the GPU against the oracle.

NFGSSolverGPU.gaussian_above_bottom_field (nfgs_gaussian_above_field; the
reference's nfgs_solver.py:4921-4943) against
OracleNFGS.gaussian_above_bottom_field (pinned in
tests/test_oracle_above_field.py) on the reference's golden boundaries, and
the analytic case of a potential field from sources above the bottom, whose
outward part is the whole observed field.
"""
import os
import sys

import numpy as np
import pytest
import torch

from oracle.nfgs_oracle import OracleNFGS
from tests.golden_cases import case_files, load, rel_l2
from tests.test_oracle_above_field import DX, DY, NX, NY, NZ, _modes, _potential_boundary

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


def _solver(nx, ny, nz, dx, dy, dt, pad, mode, rho, bx, by, bz):
    from nfgs_amd import NFGSSolverGPU
    s = NFGSSolverGPU(nx, ny, nz, dx, dy, 1.0, device="cuda",
                      dtype=torch.float64 if dt == np.float64 else torch.float32,
                      lateral_wall="paper_bc1", buffer_thickness=0, padded_buffer_zone=True,
                      pad_calculation_domain=pad, calculation_domain_mode=mode,
                      bottom_divergence_robin_weight=rho, mean_bz_policy="uniform_open",
                      horizontal_zero_mode_policy="frozen_linear_harmonic",
                      d0_zero_mode_policy="project", source_zero_mode_policy="project",
                      harmonic_residual_policy="diagnose", poloidal_update_mode="direct",
                      vertical_current_operator="staggered_face")
    if bx is not None:
        s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0,
                                     ijzbcor=3)
    return s


@pytest.mark.parametrize("name", ["mudpack_rho05", "cube16_rho05", "fp32_rho05", "unpadded_rho05",
                                  "bottompad_rho05"])
def test_above_field_matches_oracle(name):
    _announce(f"above field {name}")
    g, meta = load(CASES[name])
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    dt = np.dtype(c["dtype"])
    o = OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=dt, bottom_divergence_robin_weight=c["rho"],
                   pad_calculation_domain=c["pad"], calculation_domain_mode=c["mode"],
                   source_zero_mode_policy="project", d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"])
    s = _solver(nxo, nyo, nz, 1.0, 1.0, dt, c["pad"], c["mode"], c["rho"], g["in_bx"],
                g["in_by"], g["in_bz"])
    f = s.gaussian_above_bottom_field()
    assert tuple(f.shape) == (3, o.Ny, o.Nx)
    want = o.gaussian_above_bottom_field()
    tol = 1e-11 if dt == np.float64 else 2e-5
    got = f.double().cpu().numpy()
    for comp in range(3):
        assert rel_l2(got[comp], want[comp]) < tol, comp


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sources_above_are_all_outward(dtype):
    _announce(f"above field analytic {np.dtype(dtype).name}")
    bx, by, bz = _potential_boundary(_modes(1), -1.0)
    s = _solver(NX, NY, NZ, DX, DY, np.dtype(dtype), False, "mudpack", 0.5, bx, by, bz)
    f = s.gaussian_above_bottom_field().double().cpu().numpy()
    tol = 1e-12 if dtype == np.float64 else 2e-6
    for comp, want in enumerate((bx, by, bz)):
        assert rel_l2(f[comp], want) < tol, comp


def test_above_field_needs_a_boundary():
    _announce("above field without a boundary")
    s = _solver(NX, NY, NZ, DX, DY, np.dtype(np.float64), False, "mudpack", 0.5, None, None, None)
    with pytest.raises(RuntimeError):
        s.gaussian_above_bottom_field()
