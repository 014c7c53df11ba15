"""Two-stage poloidal update (poloidal_update_mode="two_stage") on the GPU.

nfgs_advance_ex with NFGS_POLOIDAL_TWO_STAGE against the oracle's
solve_phi_c_two_stage (nfgs_solver.py:5674-5725; pinned in
tests/test_oracle_two_stage.py) on the reference's golden boundaries, FP64 and
FP32, a tall grid whose columns take the unsegmented z-sweep, and the FP64
identity with the direct update.
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


def _gpu_solver(nxo, nyo, nz, dt, rho, pad, mode, planes, poloidal):
    from nfgs_amd import NFGSSolverGPU
    s = NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, device="cuda",
                      dtype=torch.float64 if dt == np.float64 else torch.float32,
                      pad_calculation_domain=pad, calculation_domain_mode=mode,
                      bottom_divergence_robin_weight=rho, d0_zero_mode_policy="project",
                      source_zero_mode_policy="project", poloidal_update_mode=poloidal)
    s.set_bottom_vector_boundary(*planes, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    return s


def _oracle(nxo, nyo, nz, dt, rho, pad, mode, planes, poloidal):
    o = OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=dt, bottom_divergence_robin_weight=rho,
                   pad_calculation_domain=pad, calculation_domain_mode=mode,
                   source_zero_mode_policy="project", d0_zero_mode_policy="project",
                   poloidal_update_mode=poloidal)
    o.set_bottom_vector_boundary(*planes)
    return o


def _compare(s, o, dt, iters=4):
    kw = dict(max_iter=iters, tol=1e-30, patience=50, sigma_improvement_tol=None)
    h = s.iterate(verbose=False, **kw)
    ho = o.iterate(**kw)
    tol = 1e-10 if dt == np.float64 else 2e-4
    assert rel_l2(np.array(h["sigma"]), np.array(ho["sigma"])) < tol
    assert rel_l2(np.array(h["energy"]), np.array(ho["energy"])) < tol
    f, fo = s.fields_to_numpy(), o.fields_to_numpy()
    for k in ("Bx", "By", "Bz"):
        assert rel_l2(f[k], fo[k]) < (1e-10 if dt == np.float64 else 1e-4), k
    return f


@pytest.mark.parametrize("name", ["mudpack_rho05", "mudpack_rho0", "mudpack_rho1", "fp32_rho05"])
def test_two_stage_matches_oracle(name):
    _announce(f"two-stage {name}")
    g, meta = load(CASES[name])
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    dt = np.dtype(c["dtype"])
    planes = (g["in_bx"], g["in_by"], g["in_bz"])
    args = (nxo, nyo, nz, dt, c["rho"], c["pad"], c["mode"], planes, "two_stage")
    _compare(_gpu_solver(*args), _oracle(*args), dt)


def test_two_stage_equals_direct_in_fp64():
    _announce("two-stage vs direct")
    g, meta = load(CASES["cube16_rho05"])
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    planes = (g["in_bx"], g["in_by"], g["in_bz"])
    out = {}
    for mode in ("direct", "two_stage"):
        s = _gpu_solver(nxo, nyo, nz, np.dtype(np.float64), c["rho"], c["pad"], c["mode"], planes,
                        mode)
        s.iterate(max_iter=4, tol=1e-30, patience=50, sigma_improvement_tol=None, verbose=False)
        out[mode] = s.fields_to_numpy()
    for k in out["direct"]:
        assert rel_l2(out["two_stage"][k], out["direct"][k]) < 1e-10, k


def test_two_stage_on_columns_longer_than_the_segmented_sweep():
    # FP64 segments hold 32 rows and at most 8 of them: nz = 300 takes the
    # single-thread-per-mode sweep with the undivided source
    _announce("two-stage tall grid nz=300")
    rng = np.random.default_rng(5)
    nxo, nyo, nz = 10, 9, 300
    y, x = np.mgrid[0:nyo, 0:nxo]
    bz = np.exp(-((x - 3.0) ** 2 + (y - 4.0) ** 2) / 6.0) - np.exp(-((x - 7.0) ** 2 +
                                                                   (y - 5.0) ** 2) / 5.0)
    planes = (0.1 * rng.standard_normal((nyo, nxo)), 0.1 * rng.standard_normal((nyo, nxo)), bz)
    args = (nxo, nyo, nz, np.dtype(np.float64), 0.5, False, "mudpack", planes, "two_stage")
    _compare(_gpu_solver(*args), _oracle(*args), np.float64, iters=2)
