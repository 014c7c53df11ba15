"""FP32 accuracy of the HIP path (GPU).

The modal Poisson solves divide by the Laplacian eigenvalues, so rounding
noise in the lateral transforms reaches B and J amplified.  The reference
runs its FP32 transforms as FFTs (about one ulp of error); these tests hold
the HIP transforms to that level and check that an FP32 solve follows the
FP64 solve of the same boundary.
"""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from oracle import nfgs_oracle as O

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


@pytest.mark.parametrize("n", [258, 130])
def test_fp32_dct2_is_fft_accurate(n):
    """2-D DCT-I of FP32 data: error within 1.5x of numpy's FP32 FFT form
    (an FP32-accumulated GEMM of length 258 is about 6x worse)."""
    from nfgs_amd import _lib as L
    _announce(f"fp32 dct2 accuracy n={n}")
    lib = L.load()
    d = L.GridDesc(n, n, 3, 1.0, 1.0, 1.0, L.NFGS_F32, 0.0, 1.0, 1e-20, 1, 1, 0, 0, 0, 0)
    plan = C.c_void_p()
    L.check(lib.nfgs_plan_create(C.byref(d), C.byref(plan)))
    try:
        rng = np.random.default_rng(n)
        x = np.cumsum(rng.standard_normal((n, n)), axis=1).astype(np.float32)
        ref = O.dct1(O.dct1(x.astype(np.float64), -1), -2)
        npf = O.dct1(O.dct1(x, -1), -2).astype(np.float64)
        xt = torch.tensor(x, device="cuda")
        t = torch.zeros_like(xt)
        o = torch.zeros_like(xt)
        s = torch.cuda.current_stream().cuda_stream
        L.check(lib.nfgs_transform(plan, L.NFGS_DCT1, L.NFGS_FORWARD, xt.data_ptr(), 1, n, 0,
                                   t.data_ptr(), 1, n, 0, n, n, 1, s))
        L.check(lib.nfgs_transform(plan, L.NFGS_DCT1, L.NFGS_FORWARD, t.data_ptr(), n, 1, 0,
                                   o.data_ptr(), n, 1, 0, n, n, 1, s))
        torch.cuda.synchronize()
        got = o.double().cpu().numpy()
        e_gpu = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        e_fft = np.linalg.norm(npf - ref) / np.linalg.norm(ref)
        assert e_gpu < 1.5 * e_fft + 1e-9, (e_gpu, e_fft)
    finally:
        lib.nfgs_plan_destroy(plan)


def test_fp32_solve_follows_fp64():
    """128^3 multipole boundary (seed 3), rho = 0.5: the FP32 sigma_J
    trajectory stays within 2e-3 (relative) of the FP64 one over 30
    iterations.  The reference algorithm itself in FP32 (the oracle with FFT
    transforms) differs from its FP64 run by up to 1.07e-3 on this case."""
    import bench
    from nfgs_amd import NFGSSolverGPU
    from nfgs_amd.boundary import multipole_boundary
    _announce("fp32 vs fp64 trajectory 128^3")
    n = 128
    bx, by, bz, _ = multipole_boundary(n, seed=3)
    sig = {}
    for dt in (torch.float32, torch.float64):
        s = NFGSSolverGPU(n, n, n, 1.0, 1.0, 1.0, device="cuda", dtype=dt, **bench.solver_kwargs(0.5))
        s.set_bottom_vector_boundary(*(torch.tensor(p, dtype=dt, device="cuda") for p in (bx, by, bz)),
                                     zero_net_current=False, smooth_passes=0, ijzbcor=3)
        kw = dict(bench.ITER_KW)
        kw.update(patience=10 ** 6, sigma_improvement_tol=-1.0)
        h = s.iterate(max_iter=30, **kw)
        sig[dt] = np.array(h["sigma"])
    d = np.abs(sig[torch.float32] - sig[torch.float64]) / sig[torch.float64]
    assert d.max() < 2e-3, d.max()
