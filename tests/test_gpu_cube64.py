"""The CPU-runnable 64^3 configuration, solved end to end on the GPU and by
the oracle (live, on the host).

FP32/FP64 grid Nx = Ny = Nz = 64 (calculation grid 66x66x65), dx = dy = dz
= 1, bottom Bz from two seeded Gaussian poles (+1/-1, sigma 6 cells, 20
cells apart) plus N(0, 1e-3) noise, Bx and By from a constant-alpha linear
force-free field with alpha = 0.02 (nfgs_amd.boundary.dipole_boundary),
rho = 0.5, gs_cfcf start, max_iter 500, tol 1e-5, patience 40,
sigma_improvement_tol 1e-5, sigma_best.

FP64: the whole solve equals the oracle's (stop reason, best and final
iteration, trajectories and the selected fields at 1e-10; measured 1e-13 on
B, 2e-12 on J).  FP32: the same stop decisions (sigma_stall after 42
iterations, best at 2), and the fields of the selected state and of a fixed
42-step window against the FP64 oracle within the measured floors below,
and within five times the oracle's own FP32 error.  Measured on MI355X
(relative L2, GPU FP32 / oracle FP32, both against the FP64 oracle):
selected Bx 3.1e-5 / 1.2e-5, By 9.3e-6 / 8.5e-6, Bz 1.2e-5 / 1.1e-5, J up
to 2.7e-4 / 2.2e-4; window B 7.0e-5 / 1.6e-5, J 4.5e-4 / 1.9e-4.  The 1e-5
FP32 bar on B is at the reference algorithm's own FP32 error here.
"""
import sys

import numpy as np
import pytest
import torch

from nfgs_amd.boundary import dipole_boundary
from oracle.nfgs_oracle import OracleNFGS
from tests.golden_cases import rel_l2

pytestmark = pytest.mark.gpu

KW = dict(bottom_divergence_robin_weight=0.5, d0_zero_mode_policy="project",
          source_zero_mode_policy="project")
IT = dict(max_iter=500, tol=1e-5, patience=40, sigma_improvement_tol=1e-5)
WINDOW = 42
# FP32 floors against the FP64 oracle (relative L2 per field), about twice
# the measured values
FP32_SELECTED = {"B": 6e-5, "J": 5e-4}
FP32_WINDOW = {"B": 1.5e-4, "J": 1e-3}
FP32_VS_ORACLE32 = 5.0


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


@pytest.fixture(scope="module")
def planes():
    return dipole_boundary(64, seed=0)


@pytest.fixture(scope="module")
def oracle64(planes):
    o = OracleNFGS(64, 64, 64, 1.0, 1.0, 1.0, dtype=np.float64, **KW)
    o.set_bottom_vector_boundary(*planes)
    h = o.iterate(**IT)
    return h, o.fields_to_numpy(), o


def _solver(planes, dtype):
    from nfgs_amd import NFGSSolverGPU
    s = NFGSSolverGPU(64, 64, 64, 1.0, 1.0, 1.0, device="cuda", dtype=dtype, **KW)
    s.set_bottom_vector_boundary(*planes, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    assert (s.Nz, s.Ny, s.Nx) == (65, 66, 66)
    return s


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def test_fp64_full_solve_equals_the_oracle(planes, oracle64):
    _announce("cube64 fp64 full solve")
    ho, fo, _ = oracle64
    s = _solver(planes, torch.float64)
    h = s.iterate(init_mode="gs_cfcf", mu=1, stop_mode="sigma_best", verbose=False, **IT)
    assert h["stop_reason"] == ho["stop_reason"]
    assert h["best_iter"] == ho["best_iter"] and h["final_iter"] == ho["final_iter"]
    for k in ("sigma", "divB", "energy"):
        assert rel_l2(h[k], ho[k]) < 1e-10, k
    f = s.fields_to_numpy()
    for k in fo:
        assert f[k].shape == fo[k].shape == (64, 64, 64)
        assert rel_l2(f[k], fo[k]) < 1e-10, k


def test_fp32_full_solve_and_window_against_the_oracle(planes, oracle64):
    _announce("cube64 fp32 full solve and window")
    ho, fo, _ = oracle64
    s = _solver(planes, torch.float32)
    h = s.iterate(init_mode="gs_cfcf", mu=1, stop_mode="sigma_best", verbose=False, **IT)
    assert (h["stop_reason"], h["best_iter"], h["final_iter"]) == \
        (ho["stop_reason"], ho["best_iter"], ho["final_iter"])
    f = s.fields_to_numpy()
    o32 = OracleNFGS(64, 64, 64, 1.0, 1.0, 1.0, dtype=np.float32, **KW)
    o32.set_bottom_vector_boundary(*planes)
    h32 = o32.iterate(**IT)
    assert (h32["stop_reason"], h32["best_iter"]) == (ho["stop_reason"], ho["best_iter"])
    f32 = o32.fields_to_numpy()
    for k in fo:
        e, own = rel_l2(f[k], fo[k]), rel_l2(f32[k], fo[k])
        assert e < FP32_SELECTED[k[0]] and e < FP32_VS_ORACLE32 * own, (k, e, own)
    # fixed window: WINDOW steps from the gs_cfcf state, no stop rule
    st = s.compute_initial_conditions()
    o = OracleNFGS(64, 64, 64, 1.0, 1.0, 1.0, dtype=np.float64, **KW)
    o.set_bottom_vector_boundary(*planes)
    ost, ost32 = o.initial_state(), o32.initial_state()
    for _ in range(WINDOW):
        st = s.iteration_step(*st)
        ost = o.step(*ost)
        ost32 = o32.step(*ost32)
    for i, name in ((0, "B"), (1, "J")):
        e = rel_l2(st[i].double().cpu().numpy(), ost[i])
        own = rel_l2(ost32[i], ost[i])
        assert e < FP32_WINDOW[name] and e < FP32_VS_ORACLE32 * own, (name, e, own)
