"""The per-step API of the drop-in against the oracle (GPU, FP64).

iteration_step with a caller's J that is not the state's own current: the
reference forms the parallel current from B and J[2] and takes the bottom
plane of J_par from J[:, 0] (nfgs_solver.py:6690-6711), so a perturbed J
must change the step exactly as it changes the oracle's step.  And
per_step_callback: every call gets that step's own B and J, so a callback
that keeps them holds the whole trajectory, equal to the oracle's states.
"""
import sys

import numpy as np
import pytest
import torch

from tests.golden_cases import rel_l2
from tests.test_gpu_handoff import _pair

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


def _h(t):
    return t.detach().double().cpu().numpy()


@pytest.mark.parametrize("name", ["mudpack_rho05", "unpadded_rho05", "mudpack_rho1"])
def test_iteration_step_uses_the_callers_current(name):
    _announce(f"iteration_step with a perturbed J {name}")
    o, gpu, _ = _pair(name)
    s = gpu()
    B, J, phi, az, _ = s.compute_initial_conditions()
    ob, oj, ophi, oaz = o.initial_state()
    rng = np.random.default_rng(5)
    dj = 0.05 * np.std(oj) * rng.standard_normal(oj.shape)
    Jp = oj + dj
    got = s.iteration_step(B, torch.tensor(Jp, device="cuda", dtype=torch.float64), phi, az)
    want = o.step(ob, Jp, ophi, oaz)
    for k, g, w in zip(("B", "J", "Phi_c", "Az"), got[:4], want):
        assert rel_l2(_h(g), w) < 1e-10, k
    # the perturbation matters: the unperturbed step is far from it
    plain = o.step(ob, oj, ophi, oaz)
    assert rel_l2(want[0], plain[0]) > 1e-6


@pytest.mark.parametrize("name", ["mudpack_rho05", "legacy_rho05"])
def test_callback_keeps_every_steps_state(name):
    _announce(f"per_step_callback trajectory {name}")
    o, gpu, _ = _pair(name)
    s = gpu()
    kept = []
    s.iterate(init_mode="gs_cfcf", max_iter=6, tol=1e-300, mu=1, patience=0,
              sigma_improvement_tol=None, verbose=False,
              per_step_callback=lambda n, sig, B, J: kept.append((n, sig, B, J)))
    seen = []
    o.iterate(max_iter=6, tol=0.0, patience=0, sigma_improvement_tol=None,
              callback=lambda n, sig, B, J: seen.append((n, sig, B.copy(), J.copy())))
    assert [k[0] for k in kept] == [1, 2, 3, 4, 5, 6] == [k[0] for k in seen]
    for (n, sg, Bg, Jg), (_, so, Bo, Jo) in zip(kept, seen):
        assert abs(sg - so) <= 1e-10 * abs(so), n
        assert rel_l2(_h(Bg), Bo) < 1e-10, n
        assert rel_l2(_h(Jg), Jo) < 1e-10, n
