"""nfgs_field_scores_static (the entry point that allocates nothing) against
nfgs_field_scores through the C-ABI: equal sums, bit for bit, on two calls in
a row, the second on a smaller field, and the box checks of both."""
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


def _fields(shape, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    mk = lambda: torch.randn((3,) + shape, generator=g, dtype=torch.float64).cuda().contiguous()
    return mk(), mk(), mk(), torch.randn(shape, generator=g, dtype=torch.float64).cuda().contiguous()


def _both(lib, L, bref, b, j, div, box):
    nz, ny, nx = b.shape[1:]
    s = torch.cuda.current_stream().cuda_stream
    ws = torch.empty(int(lib.nfgs_scores_workspace_bytes()) // 8, device="cuda",
                     dtype=torch.float64)
    out_ws = torch.full((L.SC_COUNT,), float("nan"), device="cuda", dtype=torch.float64)
    out_st = torch.full((L.SC_COUNT,), float("nan"), device="cuda", dtype=torch.float64)
    args = (L.ptr(bref), L.ptr(b), L.ptr(j), L.ptr(div), nz, ny, nx, *box, 1e-10)
    L.check(lib.nfgs_field_scores(*args, L.ptr(ws), L.ptr(out_ws), s), "field_scores")
    L.check(lib.nfgs_field_scores_static(*args, L.ptr(out_st), s), "field_scores_static")
    torch.cuda.synchronize()
    return out_ws.cpu().numpy(), out_st.cpu().numpy()


def test_static_scores_equal_workspace_scores_on_two_calls():
    from nfgs_amd import _lib as L
    print("[gpu-case] field_scores_static two calls", flush=True)
    sys.stdout.flush()
    lib = L.load()
    assert lib.nfgs_field_scores_static_capacity() == L.SC_COUNT * 1024
    # first call: 1024 partial blocks (the full scratch); second: a smaller
    # field and box (fewer blocks), reading what the first call left behind
    for shape, box, seed in (((40, 96, 80), (0, 40, 0, 96, 0, 80), 1),
                             ((9, 13, 21), (1, 8, 2, 12, 3, 20), 2)):
        bref, b, j, div = _fields(shape, seed)
        a, st = _both(lib, L, bref, b, j, div, box)
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(st, a)
        # optional inputs left out, as nfgs_field_scores allows
        a, st = _both(lib, L, None, b, None, None, box)
        np.testing.assert_array_equal(st, a)


def test_static_scores_reject_a_bad_box():
    from nfgs_amd import _lib as L
    lib = L.load()
    b = torch.zeros((3, 4, 5, 6), device="cuda", dtype=torch.float64)
    out = torch.zeros(L.SC_COUNT, device="cuda", dtype=torch.float64)
    rc = lib.nfgs_field_scores_static(None, L.ptr(b), None, None, 4, 5, 6, 0, 5, 0, 5, 0, 6,
                                      1e-10, L.ptr(out), None)
    assert rc == L.NFGS_ERR_ARG
    assert b"outside the field" in lib.nfgs_last_error()
