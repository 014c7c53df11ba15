"""Modal z-solves through the C-ABI against the oracle (GPU).

nfgs_robin_solve and nfgs_dirichlet_solve take the lateral transforms of the
sources and return the modal solutions (nfgs_solver.py:5512-5625, :535-608;
oracle.modal_robin_from_hat and oracle.thomas_dirichlet, both pinned against
dense solves).  The column lengths cover one segment, several segments with
padding rows, and columns long enough for the plain Thomas kernels.
"""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from oracle import nfgs_oracle as O

pytestmark = pytest.mark.gpu

NX, NY = 11, 10  # 110 modes: two blocks of 64, the second one partial


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _lam_neumann(nx, ny, dx, dy):
    kx, ky = np.arange(nx), np.arange(ny)
    lx = -4.0 * np.sin(np.pi * kx / (2.0 * (nx - 1))) ** 2 / dx ** 2
    ly = -4.0 * np.sin(np.pi * ky / (2.0 * (ny - 1))) ** 2 / dy ** 2
    return ly[:, None] + lx[None, :]


def _lam_dirichlet(nx, ny, dx, dy):
    nxi, nyi = nx - 2, ny - 2
    kx, ky = np.arange(1, nxi + 1), np.arange(1, nyi + 1)
    lx = -4.0 * np.sin(np.pi * kx / (2.0 * (nxi + 1))) ** 2 / dx ** 2
    ly = -4.0 * np.sin(np.pi * ky / (2.0 * (nyi + 1))) ** 2 / dy ** 2
    return ly[:, None] + lx[None, :]


def _plan(L, lib, nz, dtype, rho, dz):
    d = L.GridDesc(NX, NY, nz, 1.0, 1.2, dz, dtype, rho, 1.0, 1e-20, 1, nz - 1, 0, NY, 0, NX)
    plan = C.c_void_p()
    L.check(lib.nfgs_plan_create(C.byref(d), C.byref(plan)))
    return plan


def _err(got, ref):
    return np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300)


# nz: one FP32 segment (5, 66), several with padding (130, 257), Thomas
# fallback (600); FP64 segments hold 32 rows
@pytest.mark.parametrize("nz", [5, 66, 130, 257, 600])
@pytest.mark.parametrize("rho", [0.0, 0.5])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_robin_solve_matches_oracle(nz, rho, dt):
    from nfgs_amd import _lib as L
    _announce(f"robin nz={nz} rho={rho} {dt}")
    lib = L.load()
    npdt = np.float32 if dt == "f32" else np.float64
    tdt = torch.float32 if dt == "f32" else torch.float64
    dz = 0.8
    plan = _plan(L, lib, nz, L.NFGS_F32 if dt == "f32" else L.NFGS_F64, rho, dz)
    try:
        rng = np.random.default_rng(nz + int(10 * rho))
        src = np.cumsum(rng.standard_normal((nz, NY, NX)), axis=0).astype(npdt)
        bh = rng.standard_normal((NY, NX)).astype(npdt)
        qh = rng.standard_normal((NY, NX)).astype(npdt)
        bh[0, 0] = 0.0
        qh[0, 0] = 0.0
        dev = lambda a: torch.tensor(a, device="cuda", dtype=tdt)
        s_t, b_t, q_t = dev(src), dev(bh), dev(qh)
        out = torch.full_like(s_t, float("nan"))
        scr = torch.zeros_like(s_t)
        npart = int(lib.nfgs_partials_count(plan))
        part = torch.zeros(npart, device="cuda", dtype=torch.float64)
        bnd = L.Boundary()
        bnd.phi_bottom_hat = b_t.data_ptr()
        bnd.q_hat = q_t.data_ptr()
        L.check(lib.nfgs_robin_solve(plan, C.byref(bnd), s_t.data_ptr(), out.data_ptr(),
                                     scr.data_ptr(), part.data_ptr(),
                                     torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        got = out.double().cpu().numpy()
        lam = _lam_neumann(NX, NY, 1.0, 1.2)
        k = np.sqrt(np.clip(-lam, 0.0, None))
        ls = lam.copy()
        ls[0, 0] = 1.0
        sh = src.astype(np.float64) / ls[None]
        ref = O.modal_robin_from_hat(sh, bh.astype(np.float64), lam, k, dz, rho,
                                     qh.astype(np.float64) if rho > 0 else None)
        e = _err(got, ref)
        if dt == "f64":
            assert e < 1e-11, e
        else:
            # the same recurrences in FP32 (the oracle on FP32 arrays)
            ref32 = O.modal_robin_from_hat(sh.astype(np.float32), bh, lam.astype(np.float32),
                                           k.astype(np.float32), np.float32(dz), rho,
                                           qh if rho > 0 else None)
            assert e < 2.0 * _err(ref32.astype(np.float64), ref) + 1e-6, e
        # zero-mode diagnostic: sum over planes of (src[:,0,0] / (4 (nx-1)(ny-1)))^2
        pblocks = npart // L.NQ
        zm = part.cpu().numpy()[L.Q_ZERO_MODE * pblocks]
        want = np.sum((src[:, 0, 0].astype(np.float64) / (4.0 * (NX - 1) * (NY - 1))) ** 2)
        assert zm == pytest.approx(want, rel=1e-12)
    finally:
        lib.nfgs_plan_destroy(plan)


@pytest.mark.parametrize("nz", [3, 5, 66, 130, 257, 600])
@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_dirichlet_solve_matches_oracle(nz, dt):
    from nfgs_amd import _lib as L
    _announce(f"dirichlet nz={nz} {dt}")
    lib = L.load()
    npdt = np.float32 if dt == "f32" else np.float64
    tdt = torch.float32 if dt == "f32" else torch.float64
    dz = 1.3
    plan = _plan(L, lib, nz, L.NFGS_F32 if dt == "f32" else L.NFGS_F64, 0.0, dz)
    try:
        nyi, nxi = NY - 2, NX - 2
        rng = np.random.default_rng(100 + nz)
        src = np.cumsum(rng.standard_normal((nz, nyi, nxi)), axis=0).astype(npdt)
        u0 = rng.standard_normal((nyi, nxi)).astype(npdt)
        dev = lambda a: torch.tensor(a, device="cuda", dtype=tdt)
        s_t, u_t = dev(src), dev(u0)
        out = torch.full_like(s_t, float("nan"))
        scr = torch.zeros_like(s_t)
        bnd = L.Boundary()
        bnd.az_bottom_hat = u_t.data_ptr()
        L.check(lib.nfgs_dirichlet_solve(plan, C.byref(bnd), s_t.data_ptr(), out.data_ptr(),
                                         scr.data_ptr(), torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        got = out.double().cpu().numpy()
        lam = _lam_dirichlet(NX, NY, 1.0, 1.2)
        rh = -src.astype(np.float64) / lam
        ref = O.thomas_dirichlet(rh, u0.astype(np.float64), np.zeros_like(lam), lam, dz)
        e = _err(got, ref)
        if dt == "f64":
            assert e < 1e-11, e
        else:
            ref32 = O.thomas_dirichlet(rh.astype(np.float32), u0, np.zeros((nyi, nxi), np.float32),
                                       lam.astype(np.float32), np.float32(dz))
            assert e < 2.0 * _err(ref32.astype(np.float64), ref) + 1e-6, e
    finally:
        lib.nfgs_plan_destroy(plan)
