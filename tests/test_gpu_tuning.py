"""Launch tuning (nfgs_plan_set_tuning) on the grids of the batched
Robin-weight scan (GPU).

The two knobs change how work is split over workgroups, not what is
computed.  On the calculation grids of the scan (130x130x129 and
386x386x129, FP32; mudpack padding of 128 and 320 observed columns) and on a
small FP64 grid, one iteration from the same gs_cfcf state is run with the
cost model's chunking and with every z-chunk count from 1 to 16, and with
each FP32 transform workgroups-per-CU value:

* the new state (Phi_c, Az), its B and J are bitwise identical;
* the monitors (sigma, divB, E, the source sum, the zero-mode sum and the
  final pass's face and cell current sums) agree to the roundoff of the
  per-chunk partial sums (relative 1e-5 in FP32, 1e-12 in FP64);
* counts whose chunks do not fit the grid are rejected, and the plan keeps
  its previous chunking.

Tuned counts launch up to six times the cost model's workgroups on these
grids; the plan's monitor partials hold every admissible launch
(nfgs_amd/csrc/launch_shape.h, checked on the host by test_launch_shape.py).
"""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MUDPACK = dict(lateral_wall="paper_bc1", pin_potential_boundaries=False, buffer_thickness=0,
               padded_buffer_zone=True, pad_calculation_domain=True,
               calculation_domain_mode="mudpack", bottom_divergence_robin_weight=0.5,
               mean_bz_policy="uniform_open", horizontal_zero_mode_policy="frozen_linear_harmonic",
               d0_zero_mode_policy="project", source_zero_mode_policy="project",
               harmonic_residual_policy="diagnose", poloidal_update_mode="direct",
               vertical_current_operator="staggered_face")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _fits(nz, count):
    c = (nz + count - 1) // count
    return count <= nz and (count == 1 or nz - (count - 1) * c >= 2)


def _one_step(s, jp0):
    """One iteration from the start state (slot 0, whose J_par is jp0) into
    slot 1; returns the new state, its B and J, and the monitors."""
    from nfgs_amd import _lib as L
    s._ws_t["jp"].copy_(jp0)
    s._advance(1, zero_source=False)
    s._metrics_host.copy_(s._metrics)
    torch.cuda.synchronize()
    m = s._metrics_host.numpy().copy()
    B, J = s._state_fields(1, face_diag=True)
    torch.cuda.synchronize()
    mf = s._metrics.cpu().numpy()
    mon = np.array([m[L.Q_SIGMA_NUM], m[L.Q_SIGMA_DEN], m[L.Q_DIV_ABS], m[L.Q_BMAG], m[L.Q_B2],
                    m[L.Q_SRC2], m[L.Q_ZERO_MODE], mf[L.Q_FACE_J2], mf[L.Q_CELL_J2]])
    return [t.clone() for t in s._slots[1]] + [B, J], mon


@pytest.mark.parametrize("observed,nz,dtype", [(128, 128, torch.float32),
                                               (320, 128, torch.float32),
                                               (56, 30, torch.float64)],
                         ids=["130x130x129-f32", "386x386x129-f32", "66x66x31-f64"])
def test_tuned_launches_compute_the_same_step(observed, nz, dtype):
    from nfgs_amd import NFGSSolverGPU
    from nfgs_amd import _lib as L
    from nfgs_amd.boundary import multipole_boundary
    bx, by, bz, _ = multipole_boundary(observed, seed=3, sigma=(8.0 * observed / 256,
                                                                20.0 * observed / 256))
    s = NFGSSolverGPU(observed, observed, nz, 1.0, 1.0, 1.0, device="cuda", dtype=dtype,
                      **MUDPACK)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    _announce(f"tuning on {s.Nz}x{s.Ny}x{s.Nx} {dtype}")
    lib, plan = s._lib, s._plan
    s._advance(0, zero_source=True)
    jp0 = s._ws_t["jp"].clone()
    ref, mref = _one_step(s, jp0)
    rtol = 1e-5 if dtype == torch.float32 else 1e-12
    tried = 0
    for count in range(1, 18):
        rc = lib.nfgs_plan_set_tuning(plan, L.TUNE_ZCHUNKS, count)
        if not _fits(s.Nz, count) or count > 16:
            assert rc != L.NFGS_OK, count
            continue
        assert rc == L.NFGS_OK, (count, lib.nfgs_last_error())
        assert lib.nfgs_plan_get_tuning(plan, L.TUNE_ZCHUNKS) == count
        _announce(f"  z-chunks {count}")
        got, mon = _one_step(s, jp0)
        for a, b in zip(got, ref):
            assert torch.equal(a, b), count
        np.testing.assert_allclose(mon, mref, rtol=rtol, atol=0, err_msg=f"z-chunks {count}")
        tried += 1
    assert tried >= 8
    # a rejected count leaves the chunking as it was
    last = lib.nfgs_plan_get_tuning(plan, L.TUNE_ZCHUNKS)
    assert lib.nfgs_plan_set_tuning(plan, L.TUNE_ZCHUNKS, 17) != L.NFGS_OK
    assert lib.nfgs_plan_get_tuning(plan, L.TUNE_ZCHUNKS) == last
    L.check(lib.nfgs_plan_set_tuning(plan, L.TUNE_ZCHUNKS, 0))
    for b in (1, 2, 3, 0):
        L.check(lib.nfgs_plan_set_tuning(plan, L.TUNE_TRANSFORM_BLOCKS_PER_CU, b))
        _announce(f"  transform blocks per CU {b}")
        got, mon = _one_step(s, jp0)
        for a, r in zip(got, ref):
            assert torch.equal(a, r), b
        np.testing.assert_array_equal(mon, mref)
