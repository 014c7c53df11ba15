"""The fused state pass against the unfused reconstruction + face pass (GPU).

nfgs_state_pass builds B plane by plane in LDS and runs the face pass on it;
its outputs must equal nfgs_reconstruct_b followed by nfgs_face_pass on the
same (Phi_c, Az), including the metric partials.  Shapes cover a single tile,
grids whose tiling is shifted (nx = 1 mod 32, ny = 1 mod 8), z-chunk
boundaries (nz > 32) and the 258x258 lateral size of the benchmark.
"""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    from nfgs_amd import _lib as L
    return L.load()


def _rel(a, b):
    a = a.double()
    b = b.double()
    return float(torch.linalg.vector_norm(a - b) / max(float(torch.linalg.vector_norm(b)), 1e-300))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(9, 12, 14), (7, 17, 33), (6, 9, 65), (40, 26, 34),
                                   (35, 258, 258)])
def test_state_pass_equals_reconstruct_then_face(dtype, shape):
    from nfgs_amd import _lib as L
    nz, ny, nx = shape
    print(f"[gpu-case] state pass {shape} {dtype}", flush=True)
    sys.stdout.flush()
    lib = L.load()
    dev = "cuda"
    desc = L.GridDesc(nx, ny, nz, 1.0, 1.0, 1.0,
                      L.NFGS_F32 if dtype == torch.float32 else L.NFGS_F64, 0.5, 0.8, 1e-20,
                      1, nz - 1, 1, ny - 1, 2, nx - 1)
    plan = C.c_void_p()
    L.check(lib.nfgs_plan_create(C.byref(desc), C.byref(plan)))
    try:
        g = torch.Generator(device="cpu").manual_seed(nx * 1000 + ny * 10 + nz)
        rnd = lambda *s: torch.randn(*s, generator=g, dtype=torch.float64).to(dev, dtype)
        N = nz * ny * nx
        b_bg = rnd(3, nz, ny, nx)
        bz_obs = rnd(ny, nx)
        jz_t = rnd(ny, nx)
        hprof = rnd(2, nz) * 0.1
        phi = rnd(nz, ny, nx)
        az = rnd(nz, ny, nx)
        bnd = L.Boundary(b_bg.data_ptr(), bz_obs.data_ptr(), jz_t.data_ptr(), None, None, None,
                         None, None, hprof.data_ptr())
        npart = int(lib.nfgs_partials_count(plan))
        s = torch.cuda.current_stream().cuda_stream
        out = {}
        for fused in (False, True):
            B = torch.zeros(3, nz, ny, nx, device=dev, dtype=dtype)
            jp = torch.zeros_like(B)
            jo = torch.zeros_like(B)
            part = torch.zeros(npart, device=dev, dtype=torch.float64)
            met = torch.zeros(L.NQ, device=dev, dtype=torch.float64)
            if fused:
                L.check(lib.nfgs_state_pass(plan, C.byref(bnd), phi.data_ptr(), az.data_ptr(),
                                            B.data_ptr(), jp.data_ptr(), jo.data_ptr(),
                                            part.data_ptr(), s))
            else:
                L.check(lib.nfgs_reconstruct_b(plan, C.byref(bnd), phi.data_ptr(), az.data_ptr(),
                                               B.data_ptr(), s))
                L.check(lib.nfgs_face_pass(plan, C.byref(bnd), B.data_ptr(), az.data_ptr(),
                                           jp.data_ptr(), jo.data_ptr(), part.data_ptr(), s))
            L.check(lib.nfgs_reduce(plan, part.data_ptr(), met.data_ptr(), s))
            torch.cuda.synchronize()
            out[fused] = (B, jp, jo, met)
        # the same arithmetic in both kernels; only the compiler's FMA
        # contraction may differ (a few ulp)
        tol = 1e-13 if dtype == torch.float64 else 2e-6
        for name, a, b in zip(("B", "J_par", "J", "metrics"), out[True], out[False]):
            assert torch.isfinite(a).all(), name
            assert _rel(a, b) < tol, (name, _rel(a, b))
        # B without the other outputs (the iteration's own call)
        jp2 = torch.zeros(3, nz, ny, nx, device=dev, dtype=dtype)
        part2 = torch.zeros(npart, device=dev, dtype=torch.float64)
        met2 = torch.zeros(L.NQ, device=dev, dtype=torch.float64)
        L.check(lib.nfgs_state_pass(plan, C.byref(bnd), phi.data_ptr(), az.data_ptr(), None,
                                    jp2.data_ptr(), None, part2.data_ptr(), s))
        L.check(lib.nfgs_reduce(plan, part2.data_ptr(), met2.data_ptr(), s))
        torch.cuda.synchronize()
        assert torch.equal(jp2, out[True][1])
        q = [L.Q_SIGMA_NUM, L.Q_SIGMA_DEN, L.Q_DIV_ABS, L.Q_BMAG, L.Q_B2]
        assert torch.equal(met2[q], out[True][3][q])
    finally:
        lib.nfgs_plan_destroy(plan)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nxo", [34, 65])
def test_narrow_last_tile_column_against_oracle(dtype, nxo):
    """Grids whose last tile column holds two grid columns (34 = 32 + 2; 65 =
    1 mod 32, shifted tiling): those tiles number their points column-major.
    A short unpadded solve against the oracle's."""
    from nfgs_amd import NFGSSolverGPU
    from oracle.nfgs_oracle import OracleNFGS
    from tests.golden_cases import rel_l2
    print(f"[gpu-case] narrow tiles nx={nxo} {np.dtype(dtype).name}", flush=True)
    sys.stdout.flush()
    nyo, nz = 12, 9
    rng = np.random.default_rng(nxo)
    y, x = np.mgrid[0:nyo, 0:nxo]
    bz = (np.exp(-((x - 0.3 * nxo) ** 2 + (y - 5) ** 2) / 18.0)
          - np.exp(-((x - 0.7 * nxo) ** 2 + (y - 6) ** 2) / 14.0))
    bx = 0.2 * np.sin(x / 5.0) + 0.05 * rng.standard_normal((nyo, nxo))
    by = 0.2 * np.cos(y / 4.0) + 0.05 * rng.standard_normal((nyo, nxo))
    o = OracleNFGS(nxo, nyo, nz, 1.0, 1.0, 1.0, dtype=dtype, bottom_divergence_robin_weight=0.5,
                   pad_calculation_domain=False, source_zero_mode_policy="project",
                   d0_zero_mode_policy="project")
    o.set_bottom_vector_boundary(bx, by, bz)
    ho = o.iterate(max_iter=4, tol=1e-30, patience=50, sigma_improvement_tol=None)
    s = NFGSSolverGPU(nxo, nyo, nz, 1.0, 1.0, 1.0, device="cuda",
                      dtype=torch.float64 if dtype == np.float64 else torch.float32,
                      pad_calculation_domain=False, bottom_divergence_robin_weight=0.5,
                      d0_zero_mode_policy="project", source_zero_mode_policy="project")
    s.set_bottom_vector_boundary(bx, by, bz, zero_net_current=False, smooth_passes=0, ijzbcor=3)
    h = s.iterate(max_iter=4, tol=1e-30, patience=50, sigma_improvement_tol=None, verbose=False)
    tol = 1e-10 if dtype == np.float64 else 2e-4
    assert rel_l2(np.array(h["sigma"]), np.array(ho["sigma"])) < tol
    assert rel_l2(np.array(h["energy"]), np.array(ho["energy"])) < tol
    f, fo = s.fields_to_numpy(), o.fields_to_numpy()
    for k in ("Bx", "By", "Bz"):
        assert rel_l2(f[k], fo[k]) < (1e-10 if dtype == np.float64 else 1e-4), k
