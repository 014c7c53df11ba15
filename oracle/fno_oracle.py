"""CPU oracle of the learned warm start's network — TEST INFRASTRUCTURE ONLY.

A NumPy FP64 statement of the cosine-basis spectral neural operator that
nfgs_amd.learned runs on the GPU (nfgs_amd/csrc/fno.hip; it maps the gs_cfcf
state's B and parallel current and rho to a correction of that current): its
forward pass, the analytic gradients of every weight, the mean-squared-error
loss and the Adam step.  The reference has no learned start (its iterate() starts only
from gs_cfcf, nfgs_solver.py:6613-6645), so there is no reference line to
follow; the network's definition is this module's and the GPU's common
contract.  Only tests/ and bench.py's CPU-baseline leg may import it.

The number of retained modes is a parameter here (the GPU keeps 16), so the
pins can run on tiny grids.

Parity pins (tests/test_oracle_fno.py):
  * truncated DCT-II     scipy.fft.dct(type=2, norm="ortho") (the first m
                         coefficients) and idct of the zero-padded spectrum
  * spectral layer       with a mode-independent mix R and m = n the layer is
                         the pointwise map R v (orthogonal basis, closed form)
  * GELU and its slope   x Phi(x) with scipy.special.erf; slope by central
                         differences
  * gradients            central finite differences of the loss in FP64 for
                         every parameter group and the input features
  * loss                 closed form on a target that differs on two nodes
  * Adam                 first step of a constant gradient moves every
                         weight by lr sign(g) (bias-corrected moments)
No function is left "parity unpinned".
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

WIDTH, FEATURES, HIDDEN, OUTPUTS, LAYERS = 32, 8, 64, 3, 4


def dct2_basis(n: int, m: int) -> np.ndarray:
    """(n, m) orthonormal DCT-II basis: B[j, k] = c_k cos(pi k (j + 1/2) / n),
    c_0 = sqrt(1/n), c_k = sqrt(2/n)."""
    j = np.arange(n, dtype=np.float64)[:, None]
    k = np.arange(m, dtype=np.float64)[None, :]
    c = np.where(k == 0, np.sqrt(1.0 / n), np.sqrt(2.0 / n))
    return c * np.cos(np.pi * k * (j + 0.5) / n)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def bf16_round(a) -> np.ndarray:
    """Round FP32 values to bfloat16 (nearest even), returned as FP32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def param_shapes(m: int):
    shapes = [("lift.w", (WIDTH, FEATURES)), ("lift.b", (WIDTH,))]
    for l in range(LAYERS):
        shapes += [(f"layer{l}.w", (WIDTH, WIDTH)), (f"layer{l}.b", (WIDTH,)),
                   (f"layer{l}.r", (m ** 3, WIDTH, WIDTH))]
    shapes += [("head1.w", (HIDDEN, WIDTH)), ("head1.b", (HIDDEN,)),
               ("head2.w", (OUTPUTS, HIDDEN)), ("head2.b", (OUTPUTS,))]
    return shapes


def features(b, jp, rho, c_b, c_j):
    """The eight input channels (8, nz, ny, nx): B c_b, J_par c_j, rho and
    z / (nz - 1)."""
    nz = b.shape[1]
    f = np.empty((FEATURES,) + b.shape[1:])
    f[0:3] = b * c_b
    f[3:6] = jp * c_j
    f[6] = rho
    f[7] = (np.arange(nz) / max(nz - 1, 1))[:, None, None]
    return f


def _spectral(v, r, bases):
    """u_hat = P_z P_y P_x v (C, M) and s = E_y E_z (R u_hat) (C, nz, ny, m)."""
    bz_, by_, bx_ = bases
    c = v.shape[0]
    ux = np.einsum("czyx,xk->czyk", v, bx_)
    uy = np.einsum("czyk,yj->czjk", ux, by_)
    u = np.einsum("czjk,zi->cijk", uy, bz_)
    uh = u.reshape(c, -1)
    w = np.einsum("moi,im->om", r, uh).reshape(u.shape)
    sz = np.einsum("cijk,zi->czjk", w, bz_)
    s = np.einsum("czjk,yj->czyk", sz, by_)
    return uh, s


def _pointwise(w, b, v):
    return np.einsum("oi,i...->o...", w, v) + b.reshape((-1,) + (1,) * (v.ndim - 1))


def forward(params, feat, bases):
    """Network output (3, nz, ny, nx) and the cache for backward."""
    bx_ = bases[2]
    cache = {"feat": feat, "v": [], "uh": [], "s": [], "z": []}
    v = _pointwise(params["lift.w"], params["lift.b"], feat)
    for l in range(LAYERS):
        cache["v"].append(v)
        uh, s = _spectral(v, params[f"layer{l}.r"], bases)
        z = _pointwise(params[f"layer{l}.w"], params[f"layer{l}.b"], v) + \
            np.einsum("czyk,xk->czyx", s, bx_)
        cache["uh"].append(uh)
        cache["s"].append(s)
        cache["z"].append(z)
        v = gelu(z) if l < LAYERS - 1 else z
    cache["v"].append(v)
    z1 = _pointwise(params["head1.w"], params["head1.b"], v)
    h = gelu(z1)
    out = _pointwise(params["head2.w"], params["head2.b"], h)
    cache["z1"], cache["h"] = z1, h
    return out, cache


def loss(out, target):
    """Mean of (out - target)^2 over every value, and its gradient."""
    d = out - target
    return float((d * d).mean()), 2.0 * d / d.size


def backward(params, cache, gout, bases):
    """Gradients of the loss for every parameter (and the input features)."""
    bz_, by_, bx_ = bases
    g = {}
    pts = lambda a: a.reshape(a.shape[0], -1)
    h, z1 = cache["h"], cache["z1"]
    g["head2.w"] = pts(gout) @ pts(h).T
    g["head2.b"] = pts(gout).sum(axis=1)
    gh = np.einsum("oi,o...->i...", params["head2.w"], gout)
    gz1 = gh * gelu_grad(z1)
    v4 = cache["v"][LAYERS]
    g["head1.w"] = pts(gz1) @ pts(v4).T
    g["head1.b"] = pts(gz1).sum(axis=1)
    gv = np.einsum("oi,o...->i...", params["head1.w"], gz1)
    for l in reversed(range(LAYERS)):
        v, z, uh = cache["v"][l], cache["z"][l], cache["uh"][l]
        gz = gv * gelu_grad(z) if l < LAYERS - 1 else gv
        g[f"layer{l}.w"] = pts(gz) @ pts(v).T
        g[f"layer{l}.b"] = pts(gz).sum(axis=1)
        gpoint = np.einsum("oi,o...->i...", params[f"layer{l}.w"], gz)
        # adjoint of s -> E_x s is P_x; of the chain E_y E_z R P_z P_y P_x
        gs = np.einsum("czyx,xk->czyk", gz, bx_)
        gw = np.einsum("czyk,yj->czjk", gs, by_)
        gw = np.einsum("czjk,zi->cijk", gw, bz_)
        gwh = gw.reshape(gw.shape[0], -1)
        r = params[f"layer{l}.r"]
        g[f"layer{l}.r"] = np.einsum("om,im->moi", gwh, uh)
        guh = np.einsum("moi,om->im", r, gwh).reshape(gw.shape)
        gsp = np.einsum("cijk,zi->czjk", guh, bz_)
        gsp = np.einsum("czjk,yj->czyk", gsp, by_)
        gv = gpoint + np.einsum("czyk,xk->czyx", gsp, bx_)
    feat = cache["feat"]
    g["lift.w"] = pts(gv) @ pts(feat).T
    g["lift.b"] = pts(gv).sum(axis=1)
    g["features"] = np.einsum("oi,o...->i...", params["lift.w"], gv)
    return g


def adam(master, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8, grad_scale=1.0):
    """One Adam step; returns the new (master, m, v)."""
    gi = grad * grad_scale
    m = beta1 * m + (1.0 - beta1) * gi
    v = beta2 * v + (1.0 - beta2) * gi * gi
    c1, c2 = 1.0 - beta1 ** step, 1.0 - beta2 ** step
    return master - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
