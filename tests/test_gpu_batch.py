"""Batched solves (nfgs_amd.batch.solve_batch) against the same solves run one
by one with iterate() (GPU).

An uneven batch: two boundaries of different grids, three Robin weights
each, FP64.  With the default launch shapes every batched solve must equal
its sequential solve exactly (same kernels, same arguments: the sigma
history, stop reason, best iteration and selected fields); with the
autotuned shapes, which change only the summation order of the monitors,
to roundoff.  The graphs are replayed many times per solve, with the host's
stop-rule work and the best-state copies between replays, and every replay's
monitors enter the compared history.  Also: a boundary shared across Robin
weights equals a boundary set from scratch.
"""
import sys

import numpy as np
import pytest
import torch

from tests.golden_cases import rel_l2

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


KW = dict(source_zero_mode_policy="project", d0_zero_mode_policy="project")
IT = dict(max_iter=40, tol=1e-5, patience=6, sigma_improvement_tol=1e-5)


def _cases(dtype):
    from nfgs_amd.batch import Case
    from nfgs_amd.boundary import multipole_boundary
    out = []
    for seed, (nx, ny, nz) in ((4, (40, 40, 24)), (5, (56, 44, 30))):
        bx, by, bz, _ = multipole_boundary(nx, ny, seed=seed, sigma=(3.0, 5.0))
        for rho in (0.0, 0.5, 1.0):
            out.append(Case(bx, by, bz, nz, rho, boundary=seed, dtype=dtype, kwargs=dict(KW)))
    return out


def _sequential(case):
    from nfgs_amd import NFGSSolverGPU
    nyo, nxo = case.bz.shape
    s = NFGSSolverGPU(nxo, nyo, case.nz, *case.spacing, device="cuda", dtype=case.dtype,
                      bottom_divergence_robin_weight=case.rho, **case.kwargs)
    s.set_bottom_vector_boundary(case.bx, case.by, case.bz, zero_net_current=False,
                                 smooth_passes=0, ijzbcor=3)
    h = s.iterate(init_mode="gs_cfcf", mu=1, stop_mode="sigma_best", verbose=False, **IT)
    return h, s


@pytest.mark.parametrize("steps,streams", [(8, 3), (5, 1)])
def test_batch_equals_sequential_solves(steps, streams):
    _announce(f"batch vs sequential, {steps} steps per graph, {streams} streams")
    from nfgs_amd.batch import solve_batch
    cases = _cases(torch.float64)
    got = solve_batch(cases, graph_steps=steps, streams=streams, tune=False, **IT)
    for case, (hb, sb) in zip(cases, got):
        hs, ss = _sequential(case)
        assert hb["sigma"] == hs["sigma"], case.rho
        assert hb["energy"] == hs["energy"] and hb["divB"] == hs["divB"]
        for k in ("stop_reason", "best_iter", "final_iter", "best_sigma"):
            assert hb[k] == hs[k], k
        assert hb["final_vs_best_B_relative"] == pytest.approx(hs["final_vs_best_B_relative"],
                                                               rel=1e-12, abs=1e-15)
        fb, fs = sb.fields_to_numpy(), ss.fields_to_numpy()
        for k in fs:
            assert np.array_equal(fb[k], fs[k]), k


def test_tuned_batch_matches_to_roundoff():
    _announce("tuned batch vs sequential")
    from nfgs_amd.batch import solve_batch
    cases = _cases(torch.float64)
    got = solve_batch(cases, graph_steps=8, streams=4, tune=True, **IT)
    for case, (hb, sb) in zip(cases, got):
        hs, ss = _sequential(case)
        n = min(len(hb["sigma"]), len(hs["sigma"]))
        assert rel_l2(hb["sigma"][:n], hs["sigma"][:n]) < 1e-12
        assert hb["stop_reason"] == hs["stop_reason"] and hb["best_iter"] == hs["best_iter"]
        fb, fs = sb.fields_to_numpy(), ss.fields_to_numpy()
        for k in fs:
            assert rel_l2(fb[k], fs[k]) < 1e-12, k


def test_shared_boundary_equals_a_fresh_setup():
    _announce("share_boundary")
    from nfgs_amd import NFGSSolverGPU
    c = _cases(torch.float64)[4]
    nyo, nxo = c.bz.shape
    a = NFGSSolverGPU(nxo, nyo, c.nz, 1.0, 1.0, 1.0, device="cuda", dtype=torch.float64,
                      bottom_divergence_robin_weight=0.0, **KW)
    a.set_bottom_vector_boundary(c.bx, c.by, c.bz)
    b = NFGSSolverGPU(nxo, nyo, c.nz, 1.0, 1.0, 1.0, device="cuda", dtype=torch.float64,
                      bottom_divergence_robin_weight=0.5, **KW)
    b.share_boundary(a)
    f = NFGSSolverGPU(nxo, nyo, c.nz, 1.0, 1.0, 1.0, device="cuda", dtype=torch.float64,
                      bottom_divergence_robin_weight=0.5, **KW)
    f.set_bottom_vector_boundary(c.bx, c.by, c.bz)
    for key in ("B_background", "Phi_above_bottom", "Az_bottom", "Jz_toroidal_boundary",
                "_phi_bottom_hat", "_q_hat", "_az_bottom_hat"):
        assert torch.equal(getattr(b, key), getattr(f, key)), key
    hb = b.iterate(verbose=False, **IT)
    hf = f.iterate(verbose=False, **IT)
    assert hb["sigma"] == hf["sigma"]
    with pytest.raises(ValueError):
        NFGSSolverGPU(nxo, nyo, c.nz + 20, 1.0, 1.0, 1.0, device="cuda", dtype=torch.float64,
                      **KW).share_boundary(a)
