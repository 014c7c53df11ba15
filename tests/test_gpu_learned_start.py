"""init_mode="learned" in the solver: the start is one iteration step from the
gs_cfcf current plus the network's correction, kept only when its sigma_J is
below the gs_cfcf state's.  A network whose output is zero reproduces the
plain solve shifted by exactly one step; a network whose start is worse
falls back to the plain solve bit for bit."""
import numpy as np
import pytest
import torch

from nfgs_amd import learned as Lm

pytestmark = pytest.mark.gpu


def _net(head_scale=0.0, head_bias=0.0):
    p = Lm.init_parameters(0)
    p["head2.w"] = p["head2.w"] * head_scale
    p["head2.b"] = np.full_like(p["head2.b"], head_bias)
    return Lm.SpectralWarmStart(params=p)


def _solve(n, seed, mode, net=None, dtype=torch.float32, max_iter=60):
    planes, rho = Lm.case(n, seed)
    s = Lm.make_solver(n, planes, rho, dtype=dtype)
    kw = dict(Lm.ITERATE_KW, max_iter=max_iter, init_mode=mode)
    if net is not None:
        kw["learned_start"] = net
    return s, s.iterate(**kw)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_zero_correction_is_the_plain_solve_one_step_ahead(dtype):
    _, hp = _solve(32, 3, "gs_cfcf", dtype=dtype)
    _, hl = _solve(32, 3, "learned", _net(), dtype=dtype)
    d = hl["init_diagnostics"]
    assert d["init_mode"] == "learned" and d["learned_start_used"]
    assert d["sigma_learned"] == hp["sigma"][0]
    k = min(len(hp["sigma"]) - 1, len(hl["sigma"]))
    assert k >= 20
    assert hl["sigma"][:k] == hp["sigma"][1:k + 1]
    assert hl["energy"][:k] == hp["energy"][1:k + 1]


class _NaNStart:
    """A start that cannot beat gs_cfcf: its corrected current is NaN, so
    its sigma_J is not finite."""

    def predict_current(self, solver, B, jp):
        jp.fill_(float("nan"))


def test_a_worse_start_falls_back_to_the_plain_solve_exactly():
    sp, hp = _solve(32, 4, "gs_cfcf")
    sl, hl = _solve(32, 4, "learned", _NaNStart())
    d = hl["init_diagnostics"]
    assert not d["learned_start_used"]
    assert hl["sigma"] == hp["sigma"]
    assert hl["best_iter"] == hp["best_iter"] and hl["stop_reason"] == hp["stop_reason"]
    assert torch.equal(sl._B, sp._B) and torch.equal(sl._J, sp._J)


def test_learned_mode_needs_a_network():
    planes, rho = Lm.case(32, 5)
    s = Lm.make_solver(32, planes, rho)
    with pytest.raises(ValueError, match="learned_start"):
        s.iterate(init_mode="learned", max_iter=2, verbose=False)


def test_a_trained_network_shortens_the_solves_it_was_trained_on():
    """Trained on eight solves of the multipole family (34 x 34 x 33), the
    network's start is kept on most of them and cuts their iterations, and
    on held-out boundaries sigma_J never ends worse than the plain solve's
    by more than 2% (a worse start is discarded)."""
    torch.manual_seed(0)
    pairs = Lm.generate_pairs(8, 32, seed0=100)
    net = Lm.SpectralWarmStart(seed=1)
    net.norm = Lm.fit_normalisation(pairs)
    hist = Lm.train(net, pairs, 60, batch=2, lr=2e-3)
    assert hist[-1] < 0.5 * hist[0]
    rows = Lm.compare_starts(net, 32, range(100, 108))
    s = Lm.summarize(rows)
    assert s["used"] >= 5
    assert s["learned_iterations_mean"] < 0.8 * s["plain_iterations_mean"]
    held = Lm.compare_starts(net, 32, range(2000, 2006))
    for r in held:
        assert r["learned_best_sigma"] <= 1.02 * r["plain_best_sigma"]
        if not r["learned_used"]:
            assert r["learned_iterations"] == r["plain_iterations"]
