"""Launch autotuner: picks the tiled stencils' z-chunk count and the FP32
transform's workgroups per CU for one grid by timing the candidates on the
device, and remembers the choice per (grid, dtype).

The two knobs (nfgs_plan_set_tuning) change how the work is split over
workgroups, not what is computed: the transform's results do not depend on
its launch shape at all, and the z-chunk count changes only the order in
which the monitor partials are summed (sigma, divB and E move in the last
bits).  The built-in choices are a cost model (z-chunks, planned for four
resident workgroups per CU) and the occupancy limit (transform); the
measured best can still differ by a few percent.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import torch

from . import _lib as L

ZCHUNK_CANDIDATES = (0, 1, 2, 3, 4, 5, 6, 8, 10, 12, 16)
TRANSFORM_BLOCKS_CANDIDATES = (0, 1, 2)

_CACHE: dict = {}
_CACHE_FILE = Path(os.environ.get("NFGS_TUNING_CACHE",
                                  Path.home() / ".cache" / "nfgs_amd" / "tuning.json"))


def _key(solver) -> str:
    return f"{solver.Nz}x{solver.Ny}x{solver.Nx}:{str(solver.dtype).split('.')[-1]}"


def _time(fn, stream, reps):
    """Average device time of fn() in microseconds (HIP events on the stream
    the work is enqueued on)."""
    fn()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps * 1e3


def measure(solver, reps: int = 5) -> dict:
    """Time every admissible candidate of both knobs on `solver`'s grid (its
    boundary must be set) and return the timings and the fastest values."""
    lib, plan = solver._lib, solver._plan
    st = torch.cuda.current_stream(solver.device)
    s = st.cuda_stream
    phi, az = solver._slots[0]
    ws = solver._ws_t
    part = L.ptr(solver._partials)
    bnd = L.C.byref(solver._bnd)
    Nz, Ny, Nx = solver.Nz, solver.Ny, solver.Nx
    pl = Ny * Nx

    def stencils():
        L.check(lib.nfgs_state_pass(plan, bnd, L.ptr(phi), L.ptr(az), None, L.ptr(ws["jp"]), None,
                                    part, s), "state pass")
        L.check(lib.nfgs_sources(plan, L.ptr(ws["jp"]), L.ptr(ws["s_bcz"]), L.ptr(ws["s_jz"]),
                                 part, s), "sources")

    def transforms():
        L.check(lib.nfgs_transform(plan, L.NFGS_DCT1, L.NFGS_FORWARD, L.ptr(ws["s_bcz"]), 1, Nx,
                                   pl, L.ptr(ws["hat0"]), 1, Nx, pl, Nx, Ny, Nz, s), "DCT x")
        L.check(lib.nfgs_transform(plan, L.NFGS_DCT1, L.NFGS_FORWARD, L.ptr(ws["hat0"]), Nx, 1,
                                   pl, L.ptr(ws["hat1"]), Nx, 1, pl, Ny, Nx, Nz, s), "DCT y")

    out = {"zchunks": {}, "transform_blocks_per_cu": {}}
    for c in ZCHUNK_CANDIDATES:
        if lib.nfgs_plan_set_tuning(plan, L.TUNE_ZCHUNKS, c) != L.NFGS_OK:
            continue
        out["zchunks"][c] = _time(stencils, st, reps)
    for b in TRANSFORM_BLOCKS_CANDIDATES:
        L.check(lib.nfgs_plan_set_tuning(plan, L.TUNE_TRANSFORM_BLOCKS_PER_CU, b), "tuning")
        out["transform_blocks_per_cu"][b] = _time(transforms, st, reps)
    best = {k: min(v, key=v.get) for k, v in out.items()}
    # the cost model and the occupancy limit stay unless a candidate beats
    # them by more than 2% (timing noise)
    for k, v in out.items():
        if v[best[k]] > 0.98 * v.get(0, float("inf")):
            best[k] = 0
    apply(solver, best)
    return {"times_us": out, "choice": best}


def apply(solver, choice: dict) -> None:
    lib, plan = solver._lib, solver._plan
    L.check(lib.nfgs_plan_set_tuning(plan, L.TUNE_ZCHUNKS, int(choice.get("zchunks", 0))),
            "tuning")
    L.check(lib.nfgs_plan_set_tuning(plan, L.TUNE_TRANSFORM_BLOCKS_PER_CU,
                                     int(choice.get("transform_blocks_per_cu", 0))), "tuning")


def _load_file():
    try:
        return json.loads(_CACHE_FILE.read_text())
    except (OSError, ValueError):
        return {}


def tuned(solver, reps: int = 5, persist: bool = False) -> dict:
    """The tuned choice for solver's grid: from memory, from the cache file,
    or measured now; applied to the solver's plan."""
    key = _key(solver)
    if key not in _CACHE:
        disk = _load_file() if persist else {}
        if key in disk:
            _CACHE[key] = disk[key]
        else:
            _CACHE[key] = measure(solver, reps)["choice"]
            if persist:
                disk[key] = _CACHE[key]
                _CACHE_FILE.parent.mkdir(parents=True, exist_ok=True)
                _CACHE_FILE.write_text(json.dumps(disk, indent=1))
    apply(solver, _CACHE[key])
    return dict(_CACHE[key])
