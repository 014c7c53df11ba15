#!/usr/bin/env python3
"""Benchmark of the plain NFGS iterate on one MI355X.

Workload: 256x256 observed multipole boundary, Nz=256 (calculation grid
258x258x257), FP32, rho in {0, 0.5, 1}; the headline value is the wall time
of set_bottom_vector_boundary + iterate() to the sigma_best stop at rho=0.5
(max_iter 500, tol 1e-5, patience 40, sigma_improvement_tol 1e-5) with the
boundary maps already in HBM.  A step is one plain iteration (sources, both
spectral Poisson solves, reconstruction, face pass and the monitor
readback); --steps K timed steps give ms_per_step.

Also reported: per-kernel times from HIP events on the launch stream with
the dominant kernel's roofline, the per-field error of the FP32 solve
against an FP64 solve of the same boundary, and the CPU oracle timed on a
bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sec-to-converge at 256^3 (warm/batched vs plain), stencil HBM GB/s, err vs FP64"
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}
INT8_PEAK_TOPS = 5000.0  # dense int8 MFMA, MI355X


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=256, help="observed lateral size and Nz")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--max-iter", type=int, default=500)
    p.add_argument("--kernel-reps", type=int, default=20)
    p.add_argument("--cpu-iters", type=int, default=3,
                   help="timed CPU oracle iterations per thread count")
    p.add_argument("--cpu-repeats", type=int, default=2,
                   help="timed repeats of the CPU oracle's 64^3 solve (after one warm-up)")
    p.add_argument("--repeats", type=int, default=5, help="timed GPU solves per rho")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 error reference")
    p.add_argument("--no-warm", action="store_true", help="skip the warm-start solves")
    p.add_argument("--no-batch", action="store_true",
                   help="skip the batched uneven solves (33 reconstructions)")
    p.add_argument("--batch-boundaries", type=int, default=3)
    p.add_argument("--batch-rhos", type=int, default=11)
    p.add_argument("--no-two-stage", action="store_true",
                   help="skip the two-stage poloidal update solve")
    p.add_argument("--no-precision", action="store_true",
                   help="skip the precision leg (512x512x256: FP32 exact and split-bf16 "
                        "against FP64)")
    p.add_argument("--precision-n", type=int, default=512)
    p.add_argument("--no-learned", action="store_true", help="skip the learned-start leg")
    p.add_argument("--no-variants", action="store_true",
                   help="skip the solves with the other closures, starts and relaxation")
    p.add_argument("--learned-pairs", type=int, default=128,
                   help="solver-generated training pairs of the learned-start leg")
    p.add_argument("--learned-epochs", type=int, default=20)
    p.add_argument("--learned-holdout", type=int, default=16)
    p.add_argument("--learned-n", type=int, default=128, help="observed size of the learned leg")
    p.add_argument("--transforms-only", action="store_true",
                   help="time one DCT-I pass along x and y at 258x258x257, 386x386x129 and "
                        "514x514x257 for FP32 exact, FP32 split-bf16 and FP64 plans, and exit")
    p.add_argument("--steps-only", action="store_true",
                   help="profiling workload: set up the 258x258x257 FP32 solve and run only "
                        "--warmup + --steps plain iterations (no solves, no legs)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run a 256 MiB device copy for counter calibration")
    return p.parse_args()


def solver_kwargs(rho):
    return dict(lateral_wall="paper_bc1", pin_potential_boundaries=False, buffer_thickness=0,
                padded_buffer_zone=True, pad_calculation_domain=True,
                calculation_domain_mode="mudpack", bottom_divergence_robin_weight=rho,
                mean_bz_policy="uniform_open", horizontal_zero_mode_policy="frozen_linear_harmonic",
                d0_zero_mode_policy="project", source_zero_mode_policy="project",
                harmonic_residual_policy="diagnose", poloidal_update_mode="direct",
                vertical_current_operator="staggered_face")


ITER_KW = dict(init_mode="gs_cfcf", tol=1e-5, mu=1, patience=40, sigma_improvement_tol=1e-5,
               stop_mode="sigma_best", verbose=False)


def solve(torch, NFGSSolverGPU, planes, n, rho, dtype, max_iter, start=None, extra=None,
          nz=None, iterate_kw=None):
    """One timed solve with the boundary already resident on the device
    (n x n observed columns over nz planes, default n).
    start=(B, Az): init_mode="field" from that state instead of gs_cfcf;
    extra: more constructor keywords; iterate_kw: more iterate() keywords."""
    bx, by, bz = (p.to(dtype) for p in planes)
    kw = solver_kwargs(rho)
    kw.update(extra or {})
    s = NFGSSolverGPU(n, n, n if nz is None else nz, 1.0, 1.0, 1.0, device="cuda", dtype=dtype,
                      **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kw = dict(ITER_KW)
    if start is not None:
        kw.update(init_mode="field", init_field=start[0], init_az=start[1])
    kw.update(iterate_kw or {})
    h = s.iterate(max_iter=max_iter, **kw)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return s, h, t1 - t0, t2 - t0


def time_kernels(torch, s, L, reps):
    """Average device time per launch of each kernel class (HIP events on the
    stream the kernels are enqueued on)."""
    lib, plan, bnd, ws = s._lib, s._plan, L.C.byref(s._bnd), s._ws
    st = s._stream
    wt = s._ws_t
    phi, az = s._slots[0]
    Nz, Ny, Nx = s.Nz, s.Ny, s.Nx
    pl, nxi, nyi = Ny * Nx, Nx - 2, Ny - 2
    ip = nxi * nyi
    es = phi.element_size()
    part = L.ptr(s._partials)
    calls = {
        "state_pass": lambda: lib.nfgs_state_pass(plan, bnd, L.ptr(phi), L.ptr(az), None,
                                                  L.ptr(wt["jp"]), None, part, st),
        "face_pass": lambda: lib.nfgs_face_pass(plan, bnd, L.ptr(s._Bwork), L.ptr(az),
                                                L.ptr(wt["jp"]), None, part, st),
        "sources": lambda: lib.nfgs_sources(plan, L.ptr(wt["jp"]), L.ptr(wt["s_bcz"]),
                                            L.ptr(wt["s_jz"]), part, st),
        "reconstruct_b": lambda: lib.nfgs_reconstruct_b(plan, bnd, L.ptr(phi), L.ptr(az),
                                                        L.ptr(s._Bwork), st),
        "robin_sweep": lambda: lib.nfgs_robin_solve(plan, bnd, L.ptr(wt["hat1"]), L.ptr(wt["hat0"]),
                                                    L.ptr(wt["cscr"]), part, st),
        "dirichlet_sweep": lambda: lib.nfgs_dirichlet_solve(plan, bnd, L.ptr(wt["hat1"]),
                                                            L.ptr(wt["hat0"]), L.ptr(wt["cscr"]),
                                                            st),
        "dct_x": lambda: lib.nfgs_transform(plan, L.NFGS_DCT1, L.NFGS_FORWARD, L.ptr(wt["s_bcz"]),
                                            1, Nx, pl, L.ptr(wt["hat0"]), 1, Nx, pl, Nx, Ny, Nz, st),
        "dct_y": lambda: lib.nfgs_transform(plan, L.NFGS_DCT1, L.NFGS_FORWARD, L.ptr(wt["hat0"]),
                                            Nx, 1, pl, L.ptr(wt["hat1"]), Nx, 1, pl, Ny, Nx, Nz, st),
        "dst_x": lambda: lib.nfgs_transform(plan, L.NFGS_DST1, L.NFGS_FORWARD,
                                            L.ptr(wt["s_jz"]) + ip * es, 1, nxi, ip,
                                            L.ptr(wt["hat0"]) + ip * es, 1, nxi, ip, nxi, nyi,
                                            Nz - 2, st),
        "dst_y": lambda: lib.nfgs_transform(plan, L.NFGS_DST1, L.NFGS_FORWARD,
                                            L.ptr(wt["hat0"]) + ip * es, nxi, 1, ip,
                                            L.ptr(wt["hat1"]) + ip * es, nxi, 1, ip, nyi, nxi,
                                            Nz - 2, st),
    }
    out = {}
    for name, fn in calls.items():
        L.check(fn(), name)  # warm
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            L.check(fn(), name)
        e1.record()
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps * 1e-3  # seconds per launch
    return out


def kernel_model(s, esize):
    """Algorithmic bytes (or flops) per launch.  Units: one node of the
    258x258x257 grid; words per node follow the fused-plan table of the
    hot-path scope (FP words: 4 B at FP32)."""
    Nz, Ny, Nx = s.Nz, s.Ny, s.Nx
    N = Nz * Ny * Nx
    NI = (Nz - 2) * (Ny - 2) * (Nx - 2)
    return {
        # words per node: reads + writes
        # state_pass: Phi_c, Az, B_bg (3) read; J_par (3) written
        "state_pass": dict(bytes=(1 + 1 + 3 + 3) * N * esize, words=8),
        "face_pass": dict(bytes=(3 + 3 + 1 + 3) * N * esize, words=10),
        "sources": dict(bytes=(3 + 1) * N * esize + NI * esize, words=5),
        "reconstruct_b": dict(bytes=(1 + 1 + 3 + 3) * N * esize, words=8),
        "robin_sweep": dict(bytes=2 * N * esize, words=2),
        "dirichlet_sweep": dict(bytes=2 * NI * esize, words=2),
        "dct_x": dict(bytes=2 * N * esize, flops=2.0 * Nx * N,
                      int8_ops=int8_ops(Nx, Ny * Nz)),
        "dct_y": dict(bytes=2 * N * esize, flops=2.0 * Ny * N,
                      int8_ops=int8_ops(Ny, Nx * Nz)),
        "dst_x": dict(bytes=2 * NI * esize, flops=2.0 * (Nx - 2) * NI,
                      int8_ops=int8_ops(Nx - 2, (Ny - 2) * (Nz - 2))),
        "dst_y": dict(bytes=2 * NI * esize, flops=2.0 * (Ny - 2) * NI,
                      int8_ops=int8_ops(Ny - 2, (Nx - 2) * (Nz - 2))),
    }


def int8_ops(n, nvec):
    """int8 operations the FP32 transform kernel issues for nvec vectors of
    length n: per 16-vector tile, 8 wavefronts x nch K-chunks x 2 halves x 10
    digit pairs of v_mfma_i32_16x16x64_i8 (16*16*64 multiply-adds each), with
    nch = max(1, (ceil(n/2) + 62) // 64) as the plan chooses it."""
    nch = max(1, ((n + 1) // 2 + 62) // 64)
    tiles = (nvec + 15) // 16
    return tiles * 8 * nch * 2 * 10 * 16 * 16 * 64 * 2


# launches of each kernel class in one iteration (forward and inverse passes
# of each transform run the same kernel on the same shapes).  reconstruct_b
# and face_pass are the unfused API kernels (iteration_step on a given B,
# final fields); the iteration runs the fused state_pass instead.
PER_ITER = {"state_pass": 1, "face_pass": 0, "sources": 1, "reconstruct_b": 0,
            "robin_sweep": 1, "dirichlet_sweep": 1, "dct_x": 2, "dct_y": 2, "dst_x": 2,
            "dst_y": 2}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_258x258x257_f32.json")


def measured_traffic(kernel, grid_zyx):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the calibration kernel), or
    None when no summary matches this kernel and grid."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if list(rec.get("grid_zyx", [])) != list(grid_zyx):
        return None
    k = rec.get("kernels", {}).get(kernel)
    if not k:
        return None
    return int(round(k["fetch_bytes"] + k["write_bytes"]))


LATERAL_SIZES = (96, 128, 192, 256, 320)
HEIGHTS = (64, 128, 192)


def batch_cases(nboundaries, nrhos, dtype):
    """The batched uneven workload: `nboundaries` seeded multipole
    boundaries, each with a lateral size drawn from LATERAL_SIZES and a
    height from HEIGHTS, times `nrhos` Robin weights evenly in [0, 1].  Pole
    widths scale with the lateral size (U(8, 20) cells at 256) so every
    pole keeps its 3-sigma margin on the smaller grids."""
    import numpy as np
    from nfgs_amd.batch import Case
    from nfgs_amd.boundary import multipole_boundary
    cases, grids = [], []
    for b in range(nboundaries):
        rng = np.random.default_rng(1000 + b)
        n = int(rng.choice(LATERAL_SIZES))
        nz = int(rng.choice(HEIGHTS))
        sc = n / 256.0
        bx, by, bz, _ = multipole_boundary(n, seed=b, sigma=(8.0 * sc, 20.0 * sc))
        grids.append((n, n, nz))
        for rho in np.linspace(0.0, 1.0, nrhos):
            cases.append(Case(bx, by, bz, nz, float(rho), boundary=b, dtype=dtype,
                              kwargs=solver_kwargs(float(rho))))
    return cases, grids


def batched_leg(torch, a):
    """Config-D style batch: sequential plain solves (set boundary + iterate
    per case, the reference's loop) against solve_batch (shared setup,
    graph-captured steps, concurrent streams, autotuned launches)."""
    import numpy as np
    from nfgs_amd import NFGSSolverGPU
    from nfgs_amd.batch import solve_batch
    cases, grids = batch_cases(a.batch_boundaries, a.batch_rhos, torch.float32)
    it = dict(max_iter=a.max_iter, tol=1e-5, patience=40, sigma_improvement_tol=1e-5)
    # untimed warm-up of the batched path on the smallest case of each grid
    solve_batch(cases[::a.batch_rhos], **it)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq = []
    seq_phases = {"construct_and_setup": 0.0, "iterate": 0.0}
    for c in cases:
        ta = time.perf_counter()
        nyo, nxo = c.bz.shape
        s = NFGSSolverGPU(nxo, nyo, c.nz, 1.0, 1.0, 1.0, device="cuda", dtype=c.dtype,
                          **c.kwargs)
        s.set_bottom_vector_boundary(c.bx, c.by, c.bz, zero_net_current=False,
                                     smooth_passes=0, ijzbcor=3)
        tb = time.perf_counter()
        h = s.iterate(init_mode="gs_cfcf", mu=1, stop_mode="sigma_best", verbose=False, **it)
        seq_phases["construct_and_setup"] += tb - ta
        seq_phases["iterate"] += time.perf_counter() - tb
        seq.append((h, s))  # the selected fields stay on the device until after timing
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    phases = {}
    t0 = time.perf_counter()
    res = solve_batch(cases, stats=phases, **it)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t0
    dsig, dB, same_stop = 0.0, 0.0, 0
    for (hb, sb), (hs, ss) in zip(res, seq):
        bs, bi, fi, fs = hs["best_sigma"], hs["best_iter"], hs["final_iter"], ss.fields_to_numpy()
        dsig = max(dsig, abs(hb["best_sigma"] - bs) / bs)
        fb = sb.fields_to_numpy()
        num = sum(float(np.sum((fb[k].astype(np.float64) - fs[k]) ** 2)) for k in ("Bx", "By", "Bz"))
        den = sum(float(np.sum(fs[k].astype(np.float64) ** 2)) for k in ("Bx", "By", "Bz"))
        dB = max(dB, (num / den) ** 0.5)
        same_stop += int(hb["best_iter"] == bi and hb["final_iter"] == fi)
    iters = sum(h["final_iter"] for h, _ in res)
    n = len(cases)
    return dict(solves=n, grids_observed_xyz=[list(g) for g in grids],
                grids_calc_zyx=[[s.Nz, s.Ny, s.Nx] for _, s in res[::a.batch_rhos]],
                sequential_s=round(t_seq, 3), batched_s=round(t_bat, 3),
                sequential_solves_per_hour=round(n / t_seq * 3600, 1),
                batched_solves_per_hour=round(n / t_bat * 3600, 1),
                speedup=round(t_seq / t_bat, 3), total_iterations=iters,
                max_best_sigma_rel_diff=dsig, max_B_rel_diff=dB,
                same_best_and_final_iter=f"{same_stop}/{n}", batched_phases_s=phases,
                sequential_phases_s={k: round(v, 4) for k, v in seq_phases.items()},
                note="batched: boundary setup shared across Robin weights, 8-step HIP graphs, "
                     "4 streams, autotuned z-chunks and transform blocks; tuning changes "
                     "only the monitors' summation order")


def learned_leg(torch, a):
    """Learned start at 128^3 (calculation grid 130x130x129): train the
    spectral network on solver-generated pairs of the multipole family, then
    solve held-out boundaries plain and from the learned start."""
    from nfgs_amd import learned as Lm
    t0 = time.perf_counter()
    pairs = Lm.generate_pairs(a.learned_pairs, a.learned_n, seed0=0)
    t_gen = time.perf_counter() - t0
    net = Lm.SpectralWarmStart(seed=0)
    net.norm = Lm.fit_normalisation(pairs)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = Lm.train(net, pairs, a.learned_epochs, batch=4, lr=1e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    samples = len(hist) * len(pairs)
    del pairs
    rows = Lm.compare_starts(net, a.learned_n,
                             range(1_000_000, 1_000_000 + a.learned_holdout))
    out = dict(grid_observed=a.learned_n, pairs=a.learned_pairs, epochs=len(hist),
               final_train_loss=round(hist[-1], 4), generate_s=round(t_gen, 2),
               train_s=round(t_train, 2), train_ms_per_sample=round(t_train / samples * 1e3, 2),
               parameters=net.n_params, **Lm.summarize(rows))
    del net
    return out


def field_errors(f, ref):
    """Relative L2 of every field of f against ref (fields_to_numpy dicts)."""
    import numpy as np
    return {k: float(np.linalg.norm(f[k].astype(np.float64) - ref[k]) /
                     max(np.linalg.norm(ref[k]), 1e-300)) for k in f}


def precision_leg(torch, a):
    """The precision configuration: a 512x512 observed multipole boundary
    (seed a.seed) over 256 planes (calculation grid 514x514x257), rho 0.5,
    solved in FP64 (the reference of the errors), in FP32 with the exact
    lateral arithmetic (FP64 products at this length; the int8 digit form
    covers lengths up to 258) and in FP32 with the split-bf16 x3 products.
    Per-field relative L2 and the best sigma_J deviation against FP64, and
    the reference's precision-comparison metrics (common scores)."""
    import numpy as np
    from nfgs_amd import NFGSSolverGPU
    from nfgs_amd import scores as S
    from nfgs_amd.boundary import multipole_boundary
    n = a.precision_n
    bx, by, bz, _ = multipole_boundary(n, seed=a.seed)
    planes = tuple(torch.tensor(p, dtype=torch.float64, device="cuda") for p in (bx, by, bz))
    out = {"grid_observed_xyz": [n, n, n // 2]}
    runs = {}
    for tag, dtype, extra in (("fp64", torch.float64, None),
                              ("fp32_exact", torch.float32, None),
                              ("fp32_bf16x3", torch.float32,
                               {"lateral_transform_arithmetic": "bf16x3"})):
        log(f"  precision: {tag} solve at {n}x{n}x{n // 2}")
        s, h, setup_s, total_s = solve(torch, NFGSSolverGPU, planes, n, 0.5, dtype, a.max_iter,
                                       extra=extra, nz=n // 2)
        out["grid_calc_zyx"] = [s.Nz, s.Ny, s.Nx]
        runs[tag] = (h, s.fields_to_numpy())
        out[tag] = dict(sec_to_converge=round(total_s, 3), iterations=h["final_iter"],
                        best_iter=h["best_iter"], best_sigma=h["best_sigma"],
                        stop_reason=h["stop_reason"])
        del s
        torch.cuda.empty_cache()
    h64, f64 = runs["fp64"]
    box = (slice(None),) * 3
    sj64 = S.sigma_j(f64, 1.0, box)
    for tag in ("fp32_exact", "fp32_bf16x3"):
        h, f = runs[tag]
        e = field_errors(f, f64)
        e["best_sigma_rel"] = abs(h["best_sigma"] - h64["best_sigma"]) / h64["best_sigma"]
        m = S.compute_reference_metrics(f64, f, box)
        e.update({k: m[k] for k in ("C_vec", "Eprime_n", "Eprime_m", "epsilon")})
        e["sigma_J_common_fd2"] = S.sigma_j(f, 1.0, box)
        e["sigma_J_common_fd2_fp64"] = sj64
        out[tag]["vs_fp64"] = e
    return out


def transform_timings(torch, L, reps=10):
    """Average device time (HIP events, torch's stream) of one forward DCT-I
    pass along x and along y on a random volume, per grid and arithmetic."""
    import ctypes as C
    lib = L.load()
    out = {}
    for n, nz in ((258, 257), (386, 129), (514, 257)):
        for tag, f32, arith in (("fp32_exact", True, L.LATERAL_EXACT),
                                ("fp32_bf16x3", True, L.LATERAL_BF16X3),
                                ("fp64", False, L.LATERAL_EXACT)):
            desc = L.GridDesc(n, n, nz, 1.0, 1.0, 1.0, L.NFGS_F32 if f32 else L.NFGS_F64, 0.0,
                              1.0, 1e-20, 1, 1, 0, 0, 0, 0)
            plan = C.c_void_p()
            L.check(lib.nfgs_plan_create(C.byref(desc), C.byref(plan)))
            L.check(lib.nfgs_plan_set_lateral_arithmetic(plan, arith))
            dt = torch.float32 if f32 else torch.float64
            x = torch.randn(nz, n, n, dtype=dt, device="cuda")
            y = torch.empty_like(x)
            s = torch.cuda.current_stream().cuda_stream
            pl = n * n
            for axis, (es, vs) in (("x", (1, n)), ("y", (n, 1))):
                def run():
                    L.check(lib.nfgs_transform(plan, L.NFGS_DCT1, L.NFGS_FORWARD, x.data_ptr(), es,
                                               vs, pl, y.data_ptr(), es, vs, pl, n, n, nz, s))
                run()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    run()
                e1.record()
                torch.cuda.synchronize()
                out[f"{n}x{n}x{nz} {tag} DCT {axis}"] = round(e0.elapsed_time(e1) / reps * 1e3, 1)
            lib.nfgs_plan_destroy(plan)
            del x, y
    return out


def cpu_threads():
    """Thread counts of the CPU baseline: the host's cores (at most 16, this
    process's share of a one-GPU box) and 8, the reference runner's
    (scripts/run_controlled_native.py:24)."""
    host = min(os.cpu_count() or 1, 16)
    return sorted({host, min(8, host)}, reverse=True)


def cpu_baseline(n, rho, planes_np, iters, gpu_iters, threads):
    """The CPU oracle (NumPy, its FFTs on `threads` threads through
    scipy.fft workers; the elementwise work is NumPy's single thread) on a
    bounded sample of the benchmark solve: setup and gs_cfcf start once, then
    `iters` iterations timed one by one; extrapolated to the GPU solve's
    iteration count with the median iteration time."""
    import numpy as np
    from oracle import nfgs_oracle as O

    O.set_fft_workers(threads[0])
    bx, by, bz = planes_np
    t0 = time.perf_counter()
    o = O.OracleNFGS(n, n, n, 1.0, 1.0, 1.0, dtype=np.float32, bottom_divergence_robin_weight=rho,
                     source_zero_mode_policy="project", d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(bx, by, bz)
    st = o.initial_state()
    o.sigma(st[0], st[1])
    setup_init = time.perf_counter() - t0
    per_thread = {}
    for th in threads:
        O.set_fft_workers(th)
        times = []
        for _ in range(iters):
            t1 = time.perf_counter()
            st = o.step(*st)
            o.sigma(st[0], st[1]), o.divB(st[0]), o.energy(st[0])
            times.append(time.perf_counter() - t1)
        per_thread[th] = times
    O.set_fft_workers(None)
    med = float(np.median(per_thread[threads[0]]))
    return dict(setup_init_s=setup_init, s_per_iter=med, per_thread=per_thread,
                extrapolated_s=setup_init + med * gpu_iters)


def cpu_cube64(threads, repeats):
    """The CPU oracle's whole 64^3 configuration solve (two poles, FP32,
    rho = 0.5, the sigma_best stop), 1 warm-up + `repeats` timed runs."""
    import numpy as np
    from oracle import nfgs_oracle as O
    from nfgs_amd.boundary import dipole_boundary

    planes = dipole_boundary(64, seed=0)
    O.set_fft_workers(threads)
    times, h = [], None
    for r in range(repeats + 1):
        t0 = time.perf_counter()
        o = O.OracleNFGS(64, 64, 64, 1.0, 1.0, 1.0, dtype=np.float32,
                         bottom_divergence_robin_weight=0.5, source_zero_mode_policy="project",
                         d0_zero_mode_policy="project")
        o.set_bottom_vector_boundary(*planes)
        h = o.iterate(max_iter=500, tol=1e-5, patience=40, sigma_improvement_tol=1e-5)
        if r:
            times.append(time.perf_counter() - t0)
    O.set_fft_workers(None)
    return dict(median_s=float(np.median(times)), min_s=min(times), max_s=max(times),
                repeats=repeats, iterations=h["final_iter"], best_iter=h["best_iter"],
                stop_reason=h["stop_reason"])


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (the library has no multi-GPU path); got --gpus "
              f"{a.gpus}", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from nfgs_amd import NFGSSolverGPU
    from nfgs_amd import _lib as L
    from nfgs_amd.boundary import multipole_boundary

    if a.transforms_only:
        t = transform_timings(torch, L)
        for k, v in t.items():
            log(f"{k}: {v} us")
        print(json.dumps({"workload": "lateral transform passes", "us": t}))
        return
    n = a.n
    bx, by, bz, alpha = multipole_boundary(n, seed=a.seed)
    dev = torch.device("cuda")
    planes = tuple(torch.tensor(p, dtype=torch.float64, device=dev) for p in (bx, by, bz))
    torch.cuda.synchronize()
    if a.calibrate_copy:
        # one elementwise kernel that reads 256 MiB and writes 256 MiB: the
        # scale of the FETCH_SIZE / WRITE_SIZE counters
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        log("calibration kernel: neg over 256 MiB (256 MiB read + 256 MiB written)")
        del src, dst

    # warm-up (the protocol's one untimed run): a two-iteration solve loads
    # every kernel and fills the allocator's cache, so the timed solves below
    # measure the per-solve work only
    log("warm-up solve fp32 rho=0.5, 2 iterations")
    warm = solve(torch, NFGSSolverGPU, planes, n, 0.5, torch.float32, 2)
    if a.steps_only:
        s = warm[0]
        for _ in range(a.warmup):
            s._advance(1, zero_source=False)
            s._read_metrics()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            s._advance(1, zero_source=False)
            s._read_metrics()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / max(a.steps, 1) * 1e3
        log(f"steps only: {ms:.4f} ms per step over {a.steps}")
        print(json.dumps({"workload": "steps only (profiling)", "ms_per_step": ms,
                          "steps": a.steps, "grid_zyx": [s.Nz, s.Ny, s.Nx]}))
        return
    del warm

    # headline: plain method at rho 0, 0.5 and 1 (value = rho 0.5)
    per_rho = {}
    keep = None
    for rho in (0.5, 0.0, 1.0):
        times, setups = [], []
        for r in range(max(a.repeats, 1)):
            s, h, setup_s, total_s = solve(torch, NFGSSolverGPU, planes, n, rho, torch.float32,
                                           a.max_iter)
            times.append(total_s)
            setups.append(setup_s)
            if r + 1 < a.repeats:
                del s
        per_rho[str(rho)] = dict(sec_to_converge=float(np.median(times)),
                                 sec_to_converge_min=min(times), sec_to_converge_max=max(times),
                                 repeats=len(times), setup_s=float(np.median(setups)),
                                 iterations=h["final_iter"], best_iter=h["best_iter"],
                                 best_sigma=h["best_sigma"], stop_reason=h["stop_reason"])
        log(f"solve fp32 rho={rho}: median {np.median(times):.4f} s over {len(times)} "
            f"(min {min(times):.4f}, max {max(times):.4f}), {h['final_iter']} it, "
            f"best {h['best_sigma']:.5e} ({h['stop_reason']})")
        if rho == 0.5:
            keep = (s, h)
        else:
            del s
    s, h = keep
    head = per_rho["0.5"]

    # K timed steps of the plain iteration (the loop body incl. monitor readback)
    for _ in range(a.warmup):
        s._advance(1, zero_source=False)
        s._read_metrics()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        s._advance(1, zero_source=False)
        s._read_metrics()
    torch.cuda.synchronize()
    ms_per_step = (time.perf_counter() - t0) / max(a.steps, 1) * 1e3
    log(f"step: {ms_per_step:.4f} ms")

    # per-kernel device times and the dominant kernel's roofline
    esize = 4
    ktimes = time_kernels(torch, s, L, a.kernel_reps)
    model = kernel_model(s, esize)
    share = {k: ktimes[k] * PER_ITER[k] for k in ktimes}
    kernels = {}
    for k, t in ktimes.items():
        m = model[k]
        rec = dict(us=round(t * 1e6, 2), per_iter=PER_ITER[k],
                   hbm_gbs=round(m["bytes"] / t / 1e9, 1))
        if "flops" in m:
            # the FP32-equivalent rate of the exact-integer scheme (2n flops
            # per output as a plain FP32 product would count them); not a
            # utilisation of any unit
            rec["fp32_equivalent_tflops"] = round(m["flops"] / t / 1e12, 2)
            rec["int8_tops_issued"] = round(m["int8_ops"] / t / 1e12, 1)
            rec["int8_mfma_frac"] = round(m["int8_ops"] / t / 1e12 / INT8_PEAK_TOPS, 4)
        kernels[k] = rec
    dom = max(share, key=share.get)
    stencils = [k for k in ktimes if "flops" not in model[k]]
    dom_sten = max(stencils, key=lambda k: share[k])
    def roof(k):
        m, t = model[k], ktimes[k]
        if "flops" in m:
            # issued int8 MFMA operations against the int8 peak
            ach = m["int8_ops"] / t / 1e12
            return dict(kernel=k, bound="mfma", achieved=round(ach, 1),
                        peak=INT8_PEAK_TOPS, unit="TOP/s (int8)",
                        frac=round(ach / INT8_PEAK_TOPS, 4), traffic=None,
                        avg_launch_us=round(t * 1e6, 2))
        ach = m["bytes"] / t / 1e9
        return dict(kernel=k, bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                    avg_launch_us=round(t * 1e6, 2))
    roofline = roof(dom)
    stencil_roof = roof(dom_sten)
    for r in (roofline, stencil_roof):
        r["traffic"] = measured_traffic(r["kernel"], [s.Nz, s.Ny, s.Nx])
        if r["traffic"] is not None:
            # not measured by this run: counters cannot be read beside timing
            r["traffic_source"] = ("profiles/pmc_traffic_258x258x257_f32.json (rocprofv3 "
                                   "FETCH_SIZE and WRITE_SIZE passes, calibrated)")
    log(f"dominant kernel {dom}: {roofline}")

    # warm start (init_mode="field"): the next frame of the same region (poles
    # drifted by one cell, alpha x1.05) started from this frame's selected
    # state, against its own plain gs_cfcf solve
    warm = None
    if not a.no_warm:
        log("next frame: plain and warm-started solves, fp32 rho=0.5")
        nb = multipole_boundary(n, seed=a.seed, drift=1.0, alpha_scale=1.05)
        planes2 = tuple(torch.tensor(p, dtype=torch.float64, device=dev) for p in nb[:3])
        B0, Az0, _ = s.warm_start_state()
        sp, hp, _, tp = solve(torch, NFGSSolverGPU, planes2, n, 0.5, torch.float32, a.max_iter)
        sw, hw, _, tw = solve(torch, NFGSSolverGPU, planes2, n, 0.5, torch.float32, a.max_iter,
                              start=(B0, Az0))
        fp, fw = sp.fields_to_numpy(), sw.fields_to_numpy()
        bdiff = float(np.sqrt(sum(np.sum((fw[k].astype(np.float64) - fp[k]) ** 2)
                                  for k in ("Bx", "By", "Bz")) /
                              sum(np.sum(fp[k].astype(np.float64) ** 2) for k in ("Bx", "By", "Bz"))))
        warm = dict(frame="seed %d, poles drifted 1 cell, alpha x1.05" % a.seed,
                    plain_sec_to_converge=round(tp, 4), plain_iterations=hp["final_iter"],
                    plain_best_iter=hp["best_iter"], plain_best_sigma=hp["best_sigma"],
                    warm_sec_to_converge=round(tw, 4), warm_iterations=hw["final_iter"],
                    warm_best_iter=hw["best_iter"], warm_best_sigma=hw["best_sigma"],
                    warm_start_projection_error=hw["init_diagnostics"].get(
                        "field_relative_difference"),
                    speedup=round(tp / tw, 3), B_rel_diff_warm_vs_plain=bdiff)
        log(f"  plain {tp:.3f} s ({hp['final_iter']} it, best {hp['best_sigma']:.5e}); "
            f"warm {tw:.3f} s ({hw['final_iter']} it, best {hw['best_sigma']:.5e}); "
            f"B diff {bdiff:.3e}")
        del sp, sw, B0, Az0

    # batched uneven solves: many reconstructions in one graph-captured loop
    batched = None
    if not a.no_batch:
        log(f"batched uneven solves: {a.batch_boundaries} boundaries x {a.batch_rhos} rho")
        batched = batched_leg(torch, a)
        log(f"  {batched}")

    # the two-stage poloidal update (the reference's conditioning ablation) on
    # the same boundary: its time to converge and its distance from the direct
    # solve's fields
    two = None
    if not a.no_two_stage:
        log("two-stage poloidal update, fp32 rho=0.5")
        s2, h2, _, t2 = solve(torch, NFGSSolverGPU, planes, n, 0.5, torch.float32, a.max_iter,
                              extra=dict(poloidal_update_mode="two_stage"))
        f2, fd = s2.fields_to_numpy(), s.fields_to_numpy()
        bd = float(np.sqrt(sum(np.sum((f2[k].astype(np.float64) - fd[k]) ** 2)
                               for k in ("Bx", "By", "Bz")) /
                           sum(np.sum(fd[k].astype(np.float64) ** 2) for k in ("Bx", "By", "Bz"))))
        two = dict(sec_to_converge=round(t2, 4), iterations=h2["final_iter"],
                   best_iter=h2["best_iter"], best_sigma=h2["best_sigma"],
                   stop_reason=h2["stop_reason"], B_rel_diff_vs_direct=bd)
        log(f"  two-stage {t2:.3f} s ({h2['final_iter']} it, best {h2['best_sigma']:.5e}); "
            f"B diff vs direct {bd:.3e}")
        del s2

    # the method's other forms on the same boundary: the coupled closures,
    # the soft_cfcf start, a relaxed step and adaptive mu; each one's time to
    # converge, its time per iteration and its distance from the plain solve
    variants = None
    if not a.no_variants:
        variants = {}
        fd = s.fields_to_numpy()

        def bdiff(sv):
            fv = sv.fields_to_numpy()
            return float(np.sqrt(sum(np.sum((fv[k].astype(np.float64) - fd[k]) ** 2)
                                     for k in ("Bx", "By", "Bz")) /
                                 sum(np.sum(fd[k].astype(np.float64) ** 2)
                                     for k in ("Bx", "By", "Bz"))))
        forms = {
            "coupled_k0": (dict(horizontal_zero_mode_policy="coupled_k0"), {}),
            "constrained_spline_harmonic": (
                dict(horizontal_zero_mode_policy="constrained_spline_harmonic"), {}),
            "soft_cfcf_start": ({}, dict(init_mode="soft_cfcf")),
            "mu_0.5": ({}, dict(mu=0.5)),
            "adaptive_mu": ({}, dict(adaptive_mu=True)),
        }
        for name, (ctor, it) in forms.items():
            log(f"variant {name}, fp32 rho=0.5")
            sv, hv, sv_setup, tv = solve(torch, NFGSSolverGPU, planes, n, 0.5, torch.float32,
                                         a.max_iter, extra=ctor, iterate_kw=it)
            rec = dict(sec_to_converge=round(tv, 4), iterations=len(hv["sigma"]),
                       ms_per_iteration=round((tv - sv_setup) / max(len(hv["sigma"]), 1) * 1e3,
                                              4),
                       best_iter=hv["best_iter"], best_sigma=hv["best_sigma"],
                       stop_reason=hv["stop_reason"], rollbacks=hv.get("rollbacks", 0),
                       B_rel_diff_vs_plain=bdiff(sv))
            if name == "soft_cfcf_start":
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                sv.compute_initial_conditions(init_mode="soft_cfcf")
                torch.cuda.synchronize()
                rec["start_s"] = round(time.perf_counter() - t0, 4)
            variants[name] = rec
            log(f"  {name}: {tv:.3f} s, {rec['iterations']} it, best {hv['best_sigma']:.5e} "
                f"({hv['stop_reason']}); B diff vs plain {rec['B_rel_diff_vs_plain']:.3e}")
            del sv

    # error of the FP32 solve against an FP64 solve of the same boundary
    err = None
    if not a.no_fp64:
        log("solve fp64 rho=0.5 (error reference)")
        s64, h64, _, t64 = solve(torch, NFGSSolverGPU, planes, n, 0.5, torch.float64, a.max_iter)
        f32 = s.fields_to_numpy()
        f64 = s64.fields_to_numpy()
        err = {k: float(np.linalg.norm(f32[k].astype(np.float64) - f64[k]) /
                        max(np.linalg.norm(f64[k]), 1e-300)) for k in f32}
        err["best_sigma_rel"] = abs(h["best_sigma"] - h64["best_sigma"]) / h64["best_sigma"]
        err["fp64_sec_to_converge"] = t64
        err["fp64_best_iter"] = h64["best_iter"]
        del s64
        # the reference's precision comparison in its own terms: agreement
        # metrics of the FP32 fields against the FP64 ones and sigma_J of the
        # common FD2 curl, over the observed field of view (GPU scoring; the
        # oracle's NumPy scoring timed beside it unless --no-cpu)
        from nfgs_amd import scores as S
        box = (slice(None),) * 3
        t0 = time.perf_counter()
        m = S.compute_reference_metrics(f64, f32, box)
        sj32, sj64 = S.sigma_j(f32, 1.0, box), S.sigma_j(f64, 1.0, box)
        t_gpu = time.perf_counter() - t0
        cs = {k: m[k] for k in ("C_vec", "Eprime_n", "Eprime_m", "epsilon")}
        cs.update(sigma_J_common_fd2_fp32=sj32, sigma_J_common_fd2_fp64=sj64,
                  scoring_gpu_s=round(t_gpu, 4))
        # bottom-plane re-separation: E^rel of the outward Gaussian part
        from nfgs_amd import cci
        ny_o, nx_o = f64["Bz"].shape[1:]
        helper = cci.make_bgt_helper_solver(nx=nx_o, ny=ny_o, dx=1.0, dy=1.0, dz=1.0)
        cs["Erel_Bgt_bottom"] = cci.bottom_bgt_relative_error(helper, f64, f32)
        del helper
        if not a.no_cpu:
            from oracle import nfgs_oracle as O
            t0 = time.perf_counter()
            O.reference_metrics(f64, f32, box)
            O.common_sigma_j(f32, 1.0, box)
            O.common_sigma_j(f64, 1.0, box)
            cs["scoring_cpu_oracle_s"] = round(time.perf_counter() - t0, 3)
        err["common_scores"] = cs
        log(f"  common scores FP32 vs FP64: {cs}")

    learned = None
    if not a.no_learned:
        log(f"learned start: {a.learned_pairs} pairs at {a.learned_n}^3, {a.learned_epochs} "
            f"epochs, {a.learned_holdout} held-out boundaries")
        learned = learned_leg(torch, a)
        log(f"  {learned}")

    precision = None
    if not a.no_precision:
        log(f"precision leg: {a.precision_n}^2 x {a.precision_n // 2}, FP64 / FP32 exact / "
            "FP32 split-bf16")
        precision = precision_leg(torch, a)
        log(f"  {precision}")

    cpu = None
    if not a.no_cpu:
        threads = cpu_threads()
        log(f"CPU oracle sample: setup + init + {a.cpu_iters} iterations per thread count "
            f"{threads}; 64^3 solve x {a.cpu_repeats}")
        cb = cpu_baseline(n, 0.5, (bx, by, bz), a.cpu_iters, head["iterations"], threads)
        c64 = cpu_cube64(threads[0], a.cpu_repeats)
        per = {str(th): dict(s_per_iter_median=round(float(np.median(v)), 3),
                             s_per_iter_min=round(min(v), 3), s_per_iter_max=round(max(v), 3))
               for th, v in cb["per_thread"].items()}
        cpu = dict(value=round(cb["extrapolated_s"], 2), unit="s", cores=threads[0], kind="port",
                   sample=f"NumPy oracle (FFTs on {threads[0]} threads), FP32, 258x258x257, "
                          f"rho=0.5: setup+gs_cfcf init {cb['setup_init_s']:.2f} s measured once, "
                          f"{a.cpu_iters} iterations timed at {threads} threads (median "
                          f"{cb['s_per_iter']:.2f} s each at {threads[0]}), extrapolated to the "
                          f"GPU solve's {head['iterations']} iterations",
                   s_per_iter=round(cb["s_per_iter"], 3), per_threads=per,
                   host_cpu_count=os.cpu_count(),
                   cube64_full_solve=c64)
        log(f"  CPU: {cpu}")
    line = {
        "metric": METRIC,
        "value": round(head["sec_to_converge"], 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain NFGS iterate, 256x256 multipole boundary, Nz=256 "
                               "(calculation grid 258x258x257), rho=0.5",
                   "grid_zyx": [s.Nz, s.Ny, s.Nx], "seed": a.seed, "alpha": round(alpha, 5),
                   "iterations": head["iterations"], "best_iter": head["best_iter"],
                   "best_sigma": head["best_sigma"], "stop_reason": head["stop_reason"],
                   "setup_s": round(head["setup_s"], 4), "per_rho": per_rho,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "stencil_roofline": stencil_roof,
        "kernels": kernels,
        "warm_start": warm,
        "learned_start": learned,
        "batched": batched,
        "precision": precision,
        "two_stage": two,
        "variants": variants,
        "err_vs_fp64": err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
