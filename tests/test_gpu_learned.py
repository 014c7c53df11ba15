"""The learned start's kernels against the CPU oracle (oracle/fno_oracle.py):
the truncated transforms, the forward pass layer by layer, every weight
gradient, the loss and the Adam step; then the solver-facing start and its
gs_cfcf fallback."""
import numpy as np
import pytest
import torch

from oracle import fno_oracle as fo

pytestmark = pytest.mark.gpu

GRID = (16, 17, 18)  # nz, ny, nx: the smallest grids the 16-mode network takes


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    out = {}
    for name, sh in fo.param_shapes(16):
        a = scale / np.sqrt(sh[-1]) if not name.endswith(".b") else 0.3 * scale
        out[name] = rng.uniform(-a, a, size=sh).astype(np.float32)
    return out


def _oracle_params(params):
    """What the kernels compute with: bfloat16 weights, FP32 biases."""
    return {k: (fo.bf16_round(v) if not k.endswith(".b") else v).astype(np.float64)
            for k, v in params.items()}


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _net_and_buffers(params):
    from nfgs_amd.learned import SpectralWarmStart
    net = SpectralWarmStart(device="cuda", params=params)
    return net, net._buffers(GRID, True)


def _bases():
    nz, ny, nx = GRID
    return (fo.dct2_basis(nz, 16), fo.dct2_basis(ny, 16), fo.dct2_basis(nx, 16))


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("n", [16, 130, 300])
def test_axis_transforms_match_the_oracle(axis, n):
    from nfgs_amd import _lib as L
    lib = L.load()
    rng = np.random.default_rng(n)
    outer, inner = (37, 1) if axis == "x" else ((5, 16) if axis == "y" else (3, 256))
    v = rng.normal(size=(outer, n, inner))
    B = fo.dct2_basis(n, 16)
    dv = torch.as_tensor(v, dtype=torch.float32, device="cuda")
    dB = torch.as_tensor(B, dtype=torch.float32, device="cuda")
    out = torch.empty(outer, 16, inner, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    L.check(lib.nfgs_fno_axis(L.ptr(dv), L.ptr(out), L.ptr(dB), outer, n, inner,
                              L.NFGS_FNO_PROJECT, s))
    ref = np.einsum("onj,nk->okj", v, B)
    assert _rel(out.cpu().numpy(), ref) < 2e-6
    back = torch.empty(outer, n, inner, dtype=torch.float32, device="cuda")
    L.check(lib.nfgs_fno_axis(L.ptr(out), L.ptr(back), L.ptr(dB), outer, n, inner,
                              L.NFGS_FNO_EXPAND, s))
    assert _rel(back.cpu().numpy(), np.einsum("okj,nk->onj", ref, B)) < 2e-6


def test_forward_matches_the_oracle_layer_by_layer():
    params = _params(0)
    net, b = _net_and_buffers(params)
    rng = np.random.default_rng(1)
    feat = rng.normal(size=(fo.FEATURES,) + GRID)
    b.feat.copy_(torch.as_tensor(feat.reshape(fo.FEATURES, -1), dtype=torch.float32))
    net.forward(b)
    torch.cuda.synchronize()
    ref, cache = fo.forward(_oracle_params(params), feat.astype(np.float32).astype(np.float64),
                            _bases())
    for l in range(fo.LAYERS + 1):
        got = b.v[l].cpu().numpy().reshape(cache["v"][l].shape)
        assert _rel(got, cache["v"][l]) < 2e-5, l
    assert _rel(b.h.cpu().numpy().reshape(cache["h"].shape), cache["h"]) < 2e-5
    assert _rel(b.out.cpu().numpy().reshape(ref.shape), ref) < 2e-5


def test_every_gradient_matches_the_oracle():
    params = _params(2)
    net, b = _net_and_buffers(params)
    rng = np.random.default_rng(3)
    feat = rng.normal(size=(fo.FEATURES,) + GRID)
    tgt = rng.normal(size=(fo.OUTPUTS,) + GRID)
    b.feat.copy_(torch.as_tensor(feat.reshape(fo.FEATURES, -1), dtype=torch.float32))
    dt = torch.as_tensor(tgt.reshape(fo.OUTPUTS, -1), dtype=torch.float32, device="cuda")
    net.forward(b)
    loss = net.loss_and_backward(b, dt)
    torch.cuda.synchronize()
    op = _oracle_params(params)
    out, cache = fo.forward(op, feat.astype(np.float32).astype(np.float64), _bases())
    lval, gout = fo.loss(out, tgt.astype(np.float32).astype(np.float64))
    assert float(loss.item()) == pytest.approx(lval, rel=1e-5)
    ref = fo.backward(op, cache, gout, _bases())
    got = net.parameters("grad")
    for name, _ in fo.param_shapes(16):
        assert _rel(got[name], ref[name]) < 1e-4, name


def test_gradients_accumulate_and_adam_matches_the_oracle():
    params = _params(4)
    net, b = _net_and_buffers(params)
    rng = np.random.default_rng(5)
    tgt = torch.as_tensor(rng.normal(size=(fo.OUTPUTS, b.N)), dtype=torch.float32, device="cuda")
    b.feat.copy_(torch.as_tensor(rng.normal(size=(fo.FEATURES, b.N)), dtype=torch.float32))
    net.forward(b)
    net.loss_and_backward(b, tgt)
    g1 = net.parameters("grad")
    net.forward(b)
    net.loss_and_backward(b, tgt)
    g2 = net.parameters("grad")
    for name in g1:
        np.testing.assert_allclose(g2[name], 2 * g1[name], rtol=1e-5,
                                   atol=1e-7 * np.abs(g1[name]).max())
    master0 = net.parameters("master")
    net.adam_step(1e-3, batch=2)
    torch.cuda.synchronize()
    master1 = net.parameters("master")
    weights1 = net.parameters("weights")
    zero = net.parameters("grad")
    for name in g1:
        want, _, _ = fo.adam(master0[name].astype(np.float64), 0.0, 0.0,
                             g2[name].astype(np.float64), 1, lr=1e-3, grad_scale=0.5)
        np.testing.assert_allclose(master1[name], want, rtol=0, atol=2e-7 * (1 + np.abs(want).max()))
        # the kernels' weights are the master weights rounded to bfloat16
        bf = fo.bf16_round(master1[name])
        assert np.array_equal(weights1[name], bf), name
        assert not zero[name].any()


def test_targets_features_and_start_current():
    from nfgs_amd import _lib as L
    from nfgs_amd.learned import SpectralWarmStart
    lib = L.load()
    nz, ny, nx = GRID
    n = 3 * nz * ny * nx
    rng = np.random.default_rng(6)
    f = lambda a: torch.as_tensor(a, dtype=torch.float32, device="cuda")
    jp_gs, jp, B = (f(rng.normal(size=(3,) + GRID)) for _ in range(3))
    tgt = torch.empty(n, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    L.check(lib.nfgs_fno_targets(L.NFGS_F32, L.ptr(jp_gs), L.ptr(jp), n, 0.5, L.ptr(tgt), s))
    np.testing.assert_allclose(tgt.cpu().numpy(), ((jp - jp_gs) * 0.5).cpu().numpy().ravel(),
                               atol=1e-6)
    net_out = f(rng.normal(size=n))
    start = jp_gs.clone()
    L.check(lib.nfgs_fno_apply(L.NFGS_F32, L.ptr(start), L.ptr(net_out), n, 2.0, L.ptr(start), s))
    np.testing.assert_allclose(start.cpu().numpy().ravel(),
                               (jp_gs.cpu().numpy().ravel() + 2.0 * net_out.cpu().numpy()),
                               atol=1e-5)
    net = SpectralWarmStart(device="cuda")
    b = net._buffers(GRID, False)
    net.features(b, L.NFGS_F32, B, jp_gs, 0.3, 2.0)
    torch.cuda.synchronize()
    ref = fo.features(B.cpu().numpy(), jp_gs.cpu().numpy(), 0.3, 0.5, 0.5)
    np.testing.assert_allclose(b.feat.cpu().numpy().reshape(ref.shape), ref, rtol=1e-6, atol=1e-6)
