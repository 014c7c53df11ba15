"""The benchmark configurations at full size against the oracle (GPU).

Configuration B: 256x256 observed multipole boundary (seed 0), Nz = 256,
calculation grid 258x258x257, rho in {0, 0.5, 1}; and the 128^3 multipole
case (seed 3, calculation grid 130x130x129) of the precision tests.  The
oracle ran the same stages on the CPU (tests/golden/make_full_size.py);
its volumes are held as digests (tests/digest.py: plane/row/column norms,
plane sums, random-sign sketches that estimate the relative L2 error, and
1024 sampled values), and the GPU arrays are digested the same way.

Stages: the regenerated inputs, the setup tensors, the gs_cfcf state, the
states after fixed numbers of steps (iteration_step, no stop rule), and the
sigma/divB/energy trajectory of iterate() over the same fixed length.

Bars (the worst digest measure of any component):
  FP64 against the FP64 oracle: 1e-10, or three times the FP64 floor of the
  stage where conditioning puts that floor higher.  The floor is the spread
  between two correct FP64 arithmetics of the same solve: the oracle run
  with its transforms as dense matrix products against its FFT form
  (fixtures *_dense).  At 258x258x257 it is about 6e-11 on B and 4-6e-10 on
  J after 2 to 10 steps (true relative L2 4e-11 and 2.4-3e-10), because the
  modal solves divide by eigenvalues down to ~1.5e-4.
  FP32 against the FP64 oracle: within five times the oracle's own FP32
  error on the same stage (the reference algorithm in FP32, FFT transforms),
  B of the gs_cfcf state and the first two steps within twice it, the
  sigma/divB/energy trajectory within three times it.  The north star's
  1e-5 on B in FP32 is below what the reference algorithm itself reaches
  here: its own FP32 run is 3e-4 from FP64 on B already at the gs_cfcf state
  of 258x258x257 (1.5e-3 after two steps), because the modal solves amplify
  FP32 rounding by the inverse eigenvalues.  Measured on MI355X (worst
  digest measure, GPU / oracle's own): 258^3 B 2.4e-4/3.0e-4 (init),
  1.7e-3/1.5e-3 (step 2), 3.4e-3/2.0e-3 (step 10); 130^3 B 3.4e-5/4.8e-5,
  1.5e-4/8.9e-5, 3.1e-4/8.7e-5 (step 10), 5.1e-4/1.2e-4 (step 30); the
  largest ratio is 4.7 (Az, step 30 of 130^3).  The GPU's FP32 drifts from
  FP64 up to a few times faster than the oracle's over tens of steps; the
  hardware reciprocal of the alpha projection is not the cause (IEEE
  division there gives the same errors to within their scatter).
"""
import sys

import numpy as np
import pytest
import torch

from tests import full_size as F
from tests.digest import unpack, worst

pytestmark = pytest.mark.gpu

FP64_BAR = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _worst_stage(errs):
    return {k: float(max(np.max(v) for v in m.values())) for k, m in errs.items()}


FLOOR_FACTOR = 3.0


@pytest.mark.parametrize("name", ["b256_rho05_f64", "b256_rho0_f64", "b256_rho1_f64",
                                  "c128_rho05_f64"])
def test_fp64_matches_the_oracle(name):
    _announce(f"full size {name}")
    fx = F.fixture(name)
    dense = F.fixture(name + "_dense")
    s, planes = F.gpu_solver(fx["case"], torch.float64)
    assert F.input_errors(fx, planes) < 1e-13
    n = fx["case"]["n"]
    assert (s.Nz, s.Ny, s.Nx) == tuple(fx["calc_shape_zyx"]) == (n + 1, n + 2, n + 2)
    errs = _worst_stage(F.stage_errors(s, fx, fx["case"]["steps"]))
    assert {"setup_B_background", "init_B", "step2_B", "step10_B", "step10_J"} <= set(errs)
    bad = {}
    for key, e in errs.items():
        floor = worst(unpack(dense, key), unpack(fx, key))
        if e > max(FP64_BAR, FLOOR_FACTOR * floor):
            bad[key] = (e, floor)
    assert not bad, bad
    # B itself meets the 1e-10 bar through the second step
    assert max(errs["init_B"], errs["step2_B"]) < FP64_BAR
    tr = F.traj_errors(F.trajectory(s, fx["case"]["steps"]), fx)
    own = {k: float(np.max(np.abs(dense["traj_" + k] - fx["traj_" + k]) /
                           np.abs(fx["traj_" + k]))) for k in tr}
    for k in tr:
        assert tr[k] < max(FP64_BAR, FLOOR_FACTOR * own[k]), (k, tr[k], own[k])


FP32_VS_ORACLE32 = 5.0
FP32_EARLY_B = 2.0
FP32_TRAJ = 3.0


@pytest.mark.parametrize("case,f64,f32", [("b256", "b256_rho05_f64", "b256_rho05_f32"),
                                          ("c128", "c128_rho05_f64", "c128_rho05_f32")])
def test_fp32_states_against_the_fp64_and_fp32_oracle(case, f64, f32):
    _announce(f"full size fp32 {case}")
    fx64, fx32 = F.fixture(f64), F.fixture(f32)
    s, _ = F.gpu_solver(fx64["case"], torch.float32)
    errs = _worst_stage(F.stage_errors(s, fx64, fx64["case"]["steps"]))
    for key, e in errs.items():
        own = worst(unpack(fx32, key), unpack(fx64, key))
        factor = FP32_EARLY_B if key in ("init_B", "step2_B") else FP32_VS_ORACLE32
        assert e < factor * own + 1e-7, (key, e, own)
    tr = F.trajectory(s, fx64["case"]["steps"])
    e64 = F.traj_errors(tr, fx64)
    own = {k: float(np.max(np.abs(fx32["traj_" + k] - fx64["traj_" + k]) /
                           np.abs(fx64["traj_" + k]))) for k in e64}
    for k in e64:
        assert e64[k] < FP32_TRAJ * own[k] + 1e-7, (k, e64[k], own[k])
