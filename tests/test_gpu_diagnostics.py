"""record_iteration_diagnostics=True and stop_mode="fixed_point" on the GPU.

Against the reference's golden vectors (tests/golden/diag/*.npz, see
tests/diag_cases.py for the bars): the diagnostics dict of two
iteration_step calls from the gs_cfcf state, every diagnostic list of a
30-step iterate() (FP64 and FP32, and a case whose weak-field floor and
alpha limit bite), and a fixed_point run: its stop at the reference's
iteration, the criteria flag of every step, the streak, the selected final
iterate and the definition entry of the history.  On a larger grid
(multi-block launches, 130-plane columns) the diagnostics of five steps
match the oracle's live FP64 run.
"""
import os
import sys

import numpy as np
import pytest
import torch

from tests.diag_cases import BASE_KW, CASES, check_diag, load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _solver(meta):
    from nfgs_amd import NFGSSolverGPU
    c = meta["case"]
    nxo, nyo, nz = c["observed"]
    dtype = torch.float64 if c["dtype"] == "float64" else torch.float32
    kw = dict(BASE_KW, bottom_divergence_robin_weight=c["rho"],
              record_iteration_diagnostics=True, **c["extra"])
    return NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, device="cuda", dtype=dtype, **kw)


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_iteration_step_diagnostics_match_the_reference(path):
    g, meta = load(path)
    _announce(f"iteration_step diagnostics {os.path.basename(path)}")
    fp64 = meta["case"]["dtype"] == "float64"
    s = _solver(meta)
    s.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"], zero_net_current=False,
                                 smooth_passes=0, ijzbcor=3)
    st = s.compute_initial_conditions(init_mode="gs_cfcf")
    for i, want in enumerate(meta["steps"], start=1):
        st = s.iteration_step(*st[:4], mu=1.0)
        check_diag(s._last_iteration_diagnostics, want, fp64, f"step {i}")
    assert len(s._iteration_diagnostics_history) == 2


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_iterate_records_the_reference_diagnostics(path):
    g, meta = load(path)
    _announce(f"iterate with diagnostics {os.path.basename(path)}")
    c = meta["case"]
    fp64 = c["dtype"] == "float64"
    it = dict(c["iterate"])
    s = _solver(meta)
    s.set_bottom_vector_boundary(g["in_bx"], g["in_by"], g["in_bz"], zero_net_current=False,
                                 smooth_passes=0, ijzbcor=3)
    h = s.iterate(**it)
    assert h["stop_reason"] == meta["stop_reason"]
    assert (h["best_iter"], h["final_iter"]) == (meta["best_iter"], meta["final_iter"])
    assert h["selected_state"] == meta["selected_state"]
    assert h["stop_mode"] == meta["stop_mode"]
    n = len(g["hist_sigma"])
    assert len(h["sigma"]) == n
    from nfgs_amd.solver import _IterationDiagnostics
    for k in _IterationDiagnostics.KEYS:
        assert len(h[k]) == n, k
        for i in range(n):
            check_diag({k: h[k][i]}, {k: g["hist_" + k][i]}, fp64, f"iteration {i + 1}",
                       drift=10.0)
    assert set(h["iteration_diagnostics_definition"]) == \
        set(meta["iteration_diagnostics_definition"])
    if it["stop_mode"] == "fixed_point":
        assert h["fixed_point_converged"] is True and meta["fixed_point_converged"] is True
        assert h["fixed_point_streak"] == meta["fixed_point_streak"]
        assert list(h["fixed_point_criteria_met"]) == \
            [int(v) for v in g["hist_fixed_point_criteria_met"]]
        assert h["fixed_point_tolerances"] == meta["fixed_point_tolerances"]
    f = s.fields_to_numpy()
    tol = 1e-10 if fp64 else 5e-4
    for k in ("Bx", "By", "Bz"):
        a, b = f[k].astype(np.float64), g["out_" + k].astype(np.float64)
        assert np.linalg.norm(a - b) <= tol * np.linalg.norm(b), k


def test_diagnostics_match_the_oracle_on_a_larger_grid():
    """Observed 96x80x128 (calculation grid 98x98x129 with mudpack padding):
    many workgroups per launch and columns longer than one segment.  At this
    size the FP64 states of two correct arithmetics differ by 1e-11-1e-10
    after a few steps (the modal solves divide by eigenvalues down to
    ~1e-3), and a map residual divides a state difference of ~2% of the
    state, so the bar is 1e-8 (measured 1.4e-10)."""
    _announce("diagnostics vs oracle 98x98x129")
    from nfgs_amd import NFGSSolverGPU
    from nfgs_amd.boundary import multipole_boundary
    from oracle.nfgs_oracle import OracleNFGS
    bx, by, bz, _ = multipole_boundary(96, 80, seed=11, sigma=(3.0, 8.0))
    kw = dict(BASE_KW, bottom_divergence_robin_weight=0.5, record_iteration_diagnostics=True)
    s = NFGSSolverGPU(96, 80, 128, 1.0, 1.0, 1.0, device="cuda", dtype=torch.float64, **kw)
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    o = OracleNFGS(96, 80, 128, 1.0, 1.0, 1.0, dtype=np.float64, bottom_divergence_robin_weight=0.5,
                   source_zero_mode_policy="project", d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(bx, by, bz)
    assert (s.Nz, s.Ny, s.Nx) == (o.Nz, o.Ny, o.Nx)
    h = s.iterate(max_iter=5, tol=1e-300, patience=0, sigma_improvement_tol=0.0, verbose=False)
    ho = o.iterate(max_iter=5, tol=1e-300, patience=0, sigma_improvement_tol=0.0,
                   record_diagnostics=True)
    from nfgs_amd.solver import _IterationDiagnostics
    for k in _IterationDiagnostics.KEYS:
        for i in range(5):
            check_diag({k: h[k][i]}, {k: ho[k][i]}, True, f"iteration {i + 1}", fp64_tol=1e-8)


def test_unsupported_diagnostic_combinations_raise():
    from nfgs_amd import NFGSSolverGPU
    from nfgs_amd.batch import Case, solve_batch
    with pytest.raises(NotImplementedError):
        NFGSSolverGPU(12, 10, 9, 1, 1, 1, device="cuda", record_iteration_diagnostics=True,
                      poloidal_update_mode="two_stage")
    s = NFGSSolverGPU(12, 10, 9, 1, 1, 1, device="cuda")
    with pytest.raises(ValueError, match="requires record_iteration_diagnostics"):
        s.iterate(stop_mode="fixed_point")
    b = np.zeros((10, 12))
    with pytest.raises(NotImplementedError):
        solve_batch([Case(b, b, b, 9, 0.5, kwargs=dict(record_iteration_diagnostics=True))])
