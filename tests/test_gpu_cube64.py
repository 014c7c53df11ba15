"""The CPU-runnable 64^3 configuration on the GPU against the oracle.

FP32 grid Nx = Ny = Nz = 64 (calculation grid 66x66x65), dx = dy = dz = 1,
bottom Bz from two seeded Gaussian poles (+1/-1, sigma 6 cells, 20 cells
apart) plus N(0, 1e-3) noise, Bx and By from a constant-alpha linear
force-free field with alpha = 0.02 (nfgs_amd.boundary.dipole_boundary),
rho = 0.5, gs_cfcf start: the first iterations match the oracle's, and the
full solve to the sigma_best stop (max_iter 500, tol 1e-5) ends on a lower
sigma than it started from.
"""
import sys

import numpy as np
import pytest
import torch

from nfgs_amd.boundary import dipole_boundary
from oracle.nfgs_oracle import OracleNFGS
from tests.golden_cases import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _solver(planes):
    from nfgs_amd import NFGSSolverGPU
    s = NFGSSolverGPU(64, 64, 64, 1.0, 1.0, 1.0, device="cuda", dtype=torch.float32,
                      bottom_divergence_robin_weight=0.5, d0_zero_mode_policy="project",
                      source_zero_mode_policy="project")
    s.set_bottom_vector_boundary(*planes, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    return s


def test_first_iterations_match_oracle():
    print("[gpu-case] cube64 first iterations", flush=True)
    sys.stdout.flush()
    planes = dipole_boundary(64, seed=0)
    s = _solver(planes)
    assert (s.Nz, s.Ny, s.Nx) == (65, 66, 66)
    kw = dict(max_iter=5, tol=1e-30, patience=50, sigma_improvement_tol=None)
    h = s.iterate(verbose=False, **kw)
    o = OracleNFGS(64, 64, 64, 1.0, 1.0, 1.0, dtype=np.float32, bottom_divergence_robin_weight=0.5,
                   source_zero_mode_policy="project", d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(*planes)
    ho = o.iterate(**kw)
    assert rel_l2(np.array(h["sigma"]), np.array(ho["sigma"])) < 2e-4
    assert rel_l2(np.array(h["energy"]), np.array(ho["energy"])) < 2e-4
    f, fo = s.fields_to_numpy(), o.fields_to_numpy()
    for k in ("Bx", "By", "Bz"):
        assert rel_l2(f[k], fo[k]) < 1e-4, k


def test_full_solve_reaches_the_stop_rule():
    print("[gpu-case] cube64 full solve", flush=True)
    sys.stdout.flush()
    s = _solver(dipole_boundary(64, seed=0))
    h = s.iterate(init_mode="gs_cfcf", max_iter=500, tol=1e-5, mu=1, patience=40,
                  sigma_improvement_tol=1e-5, stop_mode="sigma_best", verbose=False)
    assert h["stop_reason"] in {"sigma_stall", "tol", "sigma_improvement_tol", "max_iter"}, \
        h["stop_reason"]
    assert h["best_sigma"] < h["sigma"][0]
    f = s.fields_to_numpy()
    assert all(np.isfinite(v).all() for v in f.values())
