"""GPU side of the full-size parity checks (fixtures from
tests/golden/make_full_size.py): run the same stages as the oracle run on
the HIP path and digest them the same way."""
from __future__ import annotations

import os

import numpy as np
import torch

from tests.digest import compare, digest, unpack, worst

FULL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "full")
SETUP_KEYS = ("Jz_obs", "Az_bottom", "Jz_T", "B_background", "Phi_above_bottom", "Bz_above",
              "D0", "dphi_target")
STATE_KEYS = ("B", "J", "Phi_c", "Az")


def fixture(name: str) -> dict:
    with np.load(os.path.join(FULL_DIR, name + ".npz"), allow_pickle=False) as z:
        fx = dict(z)
    fx["case"] = eval(str(fx.pop("meta")), {"__builtins__": {}})  # a dict literal of the generator
    return fx


def _host(t):
    return t.detach().double().cpu().numpy()


def gpu_solver(case: dict, dtype):
    """Solver and boundary of a fixture case, with the boundary set."""
    import bench
    from nfgs_amd import NFGSSolverGPU
    from nfgs_amd.boundary import multipole_boundary
    n = case["n"]
    bx, by, bz, _ = multipole_boundary(n, seed=case["seed"])
    s = NFGSSolverGPU(n, n, n, 1.0, 1.0, 1.0, device="cuda", dtype=dtype,
                      **bench.solver_kwargs(case["rho"]))
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    return s, (bx, by, bz)


def setup_tensors(s) -> dict:
    return {"Jz_obs": s.Jz_obs, "Az_bottom": s.Az_bottom, "Jz_T": s.Jz_toroidal_boundary,
            "B_background": s.B_background, "Phi_above_bottom": s.Phi_above_bottom,
            "Bz_above": s.Bz_above, "D0": s.D0, "dphi_target": s._bottom_dphi_c_target}


def input_errors(fx: dict, planes) -> float:
    return max(worst(digest(p), unpack(fx, "in_" + k)) for k, p in zip(("bx", "by", "bz"), planes))


def stage_errors(s, fx: dict, steps: int, digest_steps=(2, 10, 30)) -> dict:
    """Worst digest error of every stage the fixture holds: setup tensors, the
    gs_cfcf state and the states after the digested steps (fixed-length
    stepping with iteration_step).  Returns {stage_key: {measure: per-comp}}."""
    out = {}
    for k, t in setup_tensors(s).items():
        out["setup_" + k] = compare(digest(_host(t)), unpack(fx, "setup_" + k))
    st = s.compute_initial_conditions()
    for k, t in zip(STATE_KEYS, st):
        out["init_" + k] = compare(digest(_host(t)), unpack(fx, "init_" + k))
    for step in range(1, steps + 1):
        st = s.iteration_step(*st)
        if step in digest_steps:
            for k, t in zip(STATE_KEYS, st):
                key = f"step{step}_{k}"
                if key + "/shape" in fx:
                    out[key] = compare(digest(_host(t)), unpack(fx, key))
    return out


def trajectory(s, steps: int) -> dict:
    """sigma, divB and energy over a fixed number of steps (no stop rule)."""
    h = s.iterate(init_mode="gs_cfcf", max_iter=steps, tol=1e-300, mu=1, patience=0,
                  sigma_improvement_tol=None, stop_mode="sigma_best", verbose=False)
    assert h["final_iter"] == steps, h["stop_reason"]
    return {k: np.asarray(h[k], dtype=np.float64) for k in ("sigma", "divB", "energy")}


def traj_errors(got: dict, fx: dict) -> dict:
    return {k: float(np.max(np.abs(got[k] - fx["traj_" + k]) / np.abs(fx["traj_" + k])))
            for k in ("sigma", "divB", "energy")}


def worst_of(errs: dict) -> float:
    return float(max(np.max(v) for v in errs.values()))


def free(*objs):
    for o in objs:
        del o
    torch.cuda.empty_cache()
