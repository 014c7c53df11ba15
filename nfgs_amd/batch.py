"""Many reconstructions advanced together on one GPU.

The reference runs a parameter scan (boundaries x Robin weights) as a
sequential loop of independent solves
(nfgs/_backend/scripts/run_nfgs_td_robin_weight_sweep.py:510-548).  Here
the scan runs as one batch:

* the boundary setup (Gaussian separation, bottom current) is computed once
  per boundary and shared by the solves of every Robin weight
  (NFGSSolverGPU.share_boundary): it does not depend on the weight;
* each solve's iteration is captured once as a HIP graph of K steps (the
  whole kernel sequence of K iterations, each into its own state slot, with
  the monitors of every step kept on the device), so one graph launch
  replaces K x 12 kernel launches and the host reads the monitors once per K
  steps;
* the solves' graphs run on several streams at once, so small grids, which
  leave most of the GPU idle alone, overlap;
* the launch shapes of each grid come from the autotuner (nfgs_amd.autotune).

Every solve applies the reference driver's stop rules to its own monitors
step by step (the _Driver shared with iterate()); steps computed past a stop
inside the last graph are discarded.  The kernels and their arguments are
the ones iterate() runs, so each batched solve equals its sequential solve
with the same launch tuning: the same sigma history, stop and selected
fields.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from . import autotune
from .solver import COUPLED_CLOSURES, NFGSSolverGPU, _Driver

__all__ = ["Case", "solve_batch"]


@dataclass
class Case:
    """One reconstruction: a boundary (observed Bx, By, Bz planes), the
    observed grid, the Robin weight and the constructor keywords.  Cases with
    the same `boundary` key (and grid, spacing, dtype) share their setup."""
    bx: np.ndarray
    by: np.ndarray
    bz: np.ndarray
    nz: int
    rho: float
    boundary: object = None
    spacing: tuple = (1.0, 1.0, 1.0)
    dtype: torch.dtype = torch.float32
    kwargs: dict = field(default_factory=dict)


class _Run:
    """One solve's graph-captured stepping on its stream."""

    def __init__(self, solver: NFGSSolverGPU, steps: int, stream, iterate_kw: dict):
        self.s = solver
        self.k = steps
        self.stream = stream
        self.kw = iterate_kw
        # the run's buffers belong to its stream's pool, so memory freed when
        # the run ends is reused only by work ordered after it on that stream
        with torch.cuda.stream(stream):
            self.ring = [solver._new_slot() for _ in range(steps)]
            self.best = solver._new_slot()
            self.dev_m = torch.zeros(steps, L.NQ, device=solver.device, dtype=torch.float64)
        self.host_m = torch.zeros(steps, L.NQ, dtype=torch.float64).pin_memory()
        wt = solver._ws_t
        self.ws = [L.Workspace(*(L.ptr(wt[k]) for k in ("jp", "s_bcz", "s_jz", "hat0", "hat1",
                                                        "cscr")),
                               L.ptr(solver._partials), L.ptr(self.dev_m[i]))
                   for i in range(steps)]
        self.mode = L.POLOIDAL_TWO_STAGE if solver.opts.poloidal_update_mode == "two_stage" \
            else L.POLOIDAL_DIRECT
        self.graph = None
        self.event = None
        self.n = 0
        self.drv = None
        self.done = False
        self.final = None
        self.history = None

    def _step(self, i, zero_source=False):
        phi, az = self.ring[i]
        L.check(self.s._lib.nfgs_advance_ex(self.s._plan, L.C.byref(self.s._bnd),
                                            L.C.byref(self.ws[i]), L.ptr(phi), L.ptr(az), None,
                                            int(zero_source), self.mode, self.stream.cuda_stream),
                "iteration")

    def start(self):
        """gs_cfcf start (eager), then capture the K-step graph."""
        s = self.s
        with torch.cuda.stream(self.stream):
            s._advance(0, zero_source=True)
            sigma0, divB0, E0, ratio0 = s._read_metrics()
            s._check_zero_mode(ratio0)
            s._init_diagnostics = s._gs_cfcf_diagnostics()
            for t, u in zip(self.best, s._slots[0]):
                t.copy_(u)
            kw = self.kw
            self.drv = _Driver(sigma0, E0, kw["tol"], kw["patience"], kw["sigma_improvement_tol"],
                               "sigma_best", dict(update=kw["tol"], pde=1e-8,
                                                  observable=kw["tol"], patience=10))
            # every kernel of a step has run once before the capture (the
            # sources kernel is not part of the start), so no code loading
            # or one-time query happens inside it
            wt = s._ws_t
            L.check(s._lib.nfgs_sources(s._plan, L.ptr(wt["jp"]), L.ptr(wt["s_bcz"]),
                                        L.ptr(wt["s_jz"]), L.ptr(s._partials),
                                        self.stream.cuda_stream), "sources")
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph, stream=self.stream):
                for i in range(self.k):
                    self._step(i)
        if int(kw["max_iter"]) < 1:
            self._finish()

    def launch(self):
        with torch.cuda.stream(self.stream):
            self.graph.replay()
            self.host_m.copy_(self.dev_m, non_blocking=True)
            self.event = torch.cuda.Event()
            self.event.record(self.stream)

    def collect(self, clock=None):
        """Apply the stop rules to the K monitors of the last launch."""
        t = time.perf_counter()
        self.event.synchronize()
        m = self.host_m.numpy()
        best_i = None
        for i in range(self.k):
            n = self.n + 1
            sigma, divB, energy, ratio = self.s._decode_metrics(m[i])
            self.s._check_zero_mode(ratio)
            self.drv.record(n, sigma, divB, energy)
            if self.drv.accept(n, sigma):
                best_i = i
            self.n = n
            if self.drv.check_stop(n) or n >= int(self.kw["max_iter"]):
                self.final = i
                break
        with torch.cuda.stream(self.stream):
            if best_i is not None:
                for a, b in zip(self.best, self.ring[best_i]):
                    a.copy_(b)
        if clock is not None:
            clock["collect"] += time.perf_counter() - t
        if self.final is not None:
            t = time.perf_counter()
            self._finish()
            if clock is not None:
                clock["finish"] += time.perf_counter() - t

    def _finish(self):
        s = self.s
        with torch.cuda.stream(self.stream):
            final = self.ring[self.final] if self.final is not None else self.best
            same = self.drv.best_iter == self.drv.final_iter
            self.history = s._finish(self.drv, self.best, final, same)
        # the selected state's copies read the ring and the best slot
        self.stream.synchronize()
        self.done = True
        self.graph = None
        self.ring = None
        self.best = None


def _make_solver(case: Case, device) -> NFGSSolverGPU:
    nyo, nxo = np.asarray(case.bz).shape
    kw = dict(bottom_divergence_robin_weight=case.rho)
    kw.update(case.kwargs)
    return NFGSSolverGPU(nxo, nyo, case.nz, *case.spacing, device=device, dtype=case.dtype, **kw)


def solve_batch(cases: Sequence[Case], *, max_iter=500, tol=1e-5, patience=40,
                sigma_improvement_tol: Optional[float] = 1e-5, graph_steps: int = 8,
                streams: int = 4, tune: bool = True, device="cuda", max_concurrent: int = 12,
                stats: Optional[dict] = None):
    """Solve every case; returns [(history, solver)] in the order of `cases`.

    Solves of one boundary share its setup; at most `max_concurrent` solves
    hold their graphs and state rings at once, spread over `streams`
    streams (default priority).  tune=True applies the autotuner's launch
    shapes per grid.  The stop rules are the reference driver's sigma_best
    rules with the given max_iter, tol, patience and sigma_improvement_tol.
    stats (a dict) receives the host seconds of each phase: setup (solvers,
    boundaries, tuning), start (gs_cfcf starts and graph captures), launch,
    collect (waiting for and reading monitors) and finish (selected fields).
    """
    clock = {k: 0.0 for k in ("setup", "tune", "start", "launch", "collect", "finish")}
    t_setup = time.perf_counter()
    if graph_steps < 1 or streams < 1 or max_concurrent < 1:
        raise ValueError("graph_steps, streams and max_concurrent must be >= 1")
    iterate_kw = dict(max_iter=max_iter, tol=tol, patience=patience,
                      sigma_improvement_tol=sigma_improvement_tol)
    pool = [torch.cuda.Stream(device=device) for _ in range(streams)]
    solvers = []
    roots = {}
    for c in cases:
        s = _make_solver(c, device)
        if s.record_iteration_diagnostics:
            raise NotImplementedError("solve_batch does not record per-iteration diagnostics; "
                                      "run such a case with NFGSSolverGPU.iterate()")
        if s.opts.horizontal_zero_mode_policy in COUPLED_CLOSURES:
            raise NotImplementedError("solve_batch runs the frozen horizontal harmonic only; "
                                      "run a coupled closure with NFGSSolverGPU.iterate()")
        key = (id(c.boundary) if c.boundary is not None else id(c.bz),
               (s.Nz, s.Ny, s.Nx), c.spacing, c.dtype)
        if key in roots:
            s.share_boundary(roots[key])
        else:
            s.set_bottom_vector_boundary(c.bx, c.by, c.bz, zero_net_current=False,
                                         smooth_passes=0, ijzbcor=3)
            roots[key] = s
        if tune:
            t = time.perf_counter()
            autotune.tuned(s)
            clock["tune"] += time.perf_counter() - t
        solvers.append(s)
    torch.cuda.synchronize()
    clock["setup"] = time.perf_counter() - t_setup - clock["tune"]
    results = [None] * len(solvers)
    pending = list(range(len(solvers)))
    active = []
    while pending or active:
        t = time.perf_counter()
        while pending and len(active) < max_concurrent:
            idx = pending.pop(0)
            run = _Run(solvers[idx], graph_steps, pool[len(active) % streams], iterate_kw)
            run.start()
            if run.done:
                results[idx] = (run.history, solvers[idx])
            else:
                active.append((idx, run))
        clock["start"] += time.perf_counter() - t
        t = time.perf_counter()
        for _, run in active:
            run.launch()
        clock["launch"] += time.perf_counter() - t
        still = []
        for idx, run in active:
            run.collect(clock)
            if run.done:
                results[idx] = (run.history, solvers[idx])
            else:
                still.append((idx, run))
        active = still
    torch.cuda.synchronize()
    if stats is not None:
        stats.update({k: round(v, 4) for k, v in clock.items()})
    return results
