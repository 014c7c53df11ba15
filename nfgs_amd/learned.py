"""Learned warm start for the NFGS iterate (an extension; the reference starts
only from gs_cfcf, nfgs_solver.py:6613-6645).

A small cosine-basis spectral neural operator maps what the solver has after
its gs_cfcf start (that state's B and parallel current J_par, and the Robin
weight rho) to a correction of J_par towards the converged state's.  One
solver step from the corrected current (the exact Poisson solves of the
iteration) gives the learned start, so prediction errors reach the state only
through the inverse Laplacians.  ``iterate(init_mode="learned",
learned_start=net)`` evaluates sigma_J of both starts and keeps the learned
one only when its sigma_J is lower; otherwise the solve is the plain gs_cfcf
solve.  The iteration itself is unchanged, so the converged result is judged
by the same monitors and stop rules.

Network: lift 8 -> 32, four spectral layers (truncated orthonormal DCT-II of
16 modes per axis, a 32 x 32 mix per mode, a pointwise 32 x 32 path, GELU
between layers), head 32 -> 64 -> 3.  Every pass runs in the library's
kernels (nfgs_fno_*, nfgs_amd/csrc/fno.hip): bfloat16 weights, FP32
activations and accumulation; Adam keeps FP32 master weights.  Training pairs
come from the solver itself (``generate_pairs``).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib as L

WIDTH, MODES, FEATURES, HIDDEN, OUTPUTS, LAYERS = 32, 16, 8, 64, 3, 4
NMODES = MODES ** 3


def dct2_basis(n: int, m: int = MODES) -> np.ndarray:
    """(n, m) orthonormal DCT-II basis B[j, k] = c_k cos(pi k (j + 1/2) / n)."""
    j = np.arange(n, dtype=np.float64)[:, None]
    k = np.arange(m, dtype=np.float64)[None, :]
    c = np.where(k == 0, math.sqrt(1.0 / n), math.sqrt(2.0 / n))
    return c * np.cos(np.pi * k * (j + 0.5) / n)


def param_shapes():
    shapes = [("lift.w", (WIDTH, FEATURES)), ("lift.b", (WIDTH,))]
    for l in range(LAYERS):
        shapes += [(f"layer{l}.w", (WIDTH, WIDTH)), (f"layer{l}.b", (WIDTH,)),
                   (f"layer{l}.r", (NMODES, WIDTH, WIDTH))]
    shapes += [("head1.w", (HIDDEN, WIDTH)), ("head1.b", (HIDDEN,)),
               ("head2.w", (OUTPUTS, HIDDEN)), ("head2.b", (OUTPUTS,))]
    return shapes


def init_parameters(seed: int = 0) -> dict:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for every weight and bias,
    the spectral mixes included; the last layer scaled by 0.1 so that an
    untrained network starts near the gs_cfcf state."""
    rng = np.random.default_rng(seed)
    out = {}
    for name, shape in param_shapes():
        fan_in = shape[-1] if not name.endswith(".b") else \
            dict(param_shapes())[name[:-2] + ".w"][-1]
        a = 1.0 / math.sqrt(fan_in)
        if name.startswith("head2"):
            a *= 0.1
        out[name] = rng.uniform(-a, a, size=shape).astype(np.float32)
    return out


@dataclass
class Normalisation:
    """Fixed channel scales, in units of the boundary's RMS Bz (s_B): B is
    divided by s_B * b, J_par by s_B * j and the target correction of J_par by
    s_B * out."""
    b: float = 1.0
    j: float = 1.0
    out: float = 1.0

    def as_array(self):
        return np.array([self.b, self.j, self.out])


def _bz_scale(solver) -> float:
    bz = solver.Bz_obs
    return max(float(torch.sqrt(torch.mean(bz.double() * bz.double())).item()), 1e-30)


class _Buffers:
    """Device buffers of one grid (activations kept for backward when training)."""

    def __init__(self, grid, device, train):
        nz, ny, nx = grid
        N = nz * ny * nx
        f32 = dict(device=device, dtype=torch.float32)
        self.grid, self.N = grid, N
        self.bases = [torch.as_tensor(dct2_basis(n), **f32).contiguous() for n in (nz, ny, nx)]
        self.feat = torch.empty(FEATURES, N, **f32)
        nv = LAYERS + 1 if train else 2
        self.v = [torch.empty(WIDTH, N, **f32) for _ in range(nv)]
        self.ux = torch.empty(WIDTH * nz * ny * MODES, **f32)
        self.uy = torch.empty(WIDTH * nz * MODES * MODES, **f32)
        self.uh = [torch.empty(WIDTH, NMODES, **f32) for _ in range(LAYERS if train else 1)]
        self.w = torch.empty(WIDTH, NMODES, **f32)
        self.sz = torch.empty(WIDTH * nz * MODES * MODES, **f32)
        self.spec = [torch.empty(WIDTH * nz * ny * MODES, **f32) for _ in range(LAYERS if train
                                                                               else 1)]
        self.h = torch.empty(HIDDEN, N, **f32)
        self.out = torch.empty(OUTPUTS, N, **f32)
        self.train = train
        if train:
            self.gout = torch.empty(OUTPUTS, N, **f32)
            self.gh = torch.empty(HIDDEN, N, **f32)
            self.gz1 = torch.empty(HIDDEN, N, **f32)
            self.P = [torch.empty(WIDTH, N, **f32) for _ in range(2)]
            self.gz = torch.empty(WIDTH, N, **f32)
            self.gspec = [torch.empty(WIDTH * nz * ny * MODES, **f32) for _ in range(2)]
            self.gw = torch.empty(WIDTH, NMODES, **f32)
            self.guh = torch.empty(WIDTH, NMODES, **f32)
            ws = max(int(L.load().nfgs_fno_outer_sum_workspace_bytes(a, c))
                     for a, c in ((WIDTH, FEATURES), (WIDTH, WIDTH), (HIDDEN, WIDTH),
                                  (OUTPUTS, HIDDEN)))
            self.os_ws = torch.empty(ws, device=device, dtype=torch.uint8)
            self.loss_ws = torch.empty(int(L.load().nfgs_fno_loss_workspace_bytes()),
                                       device=device, dtype=torch.uint8)
            self.loss = torch.zeros(1, device=device, dtype=torch.float64)


class SpectralWarmStart:
    """The learned start's network: parameters, forward and backward passes
    and the Adam step, all on the device through the library's kernels."""

    def __init__(self, device="cuda", seed: int = 0, params: Optional[dict] = None,
                 normalisation: Optional[Normalisation] = None):
        if not torch.cuda.is_available():
            raise RuntimeError("the learned start runs on a ROCm GPU only")
        self.device = torch.device(device)
        self._lib = L.load()
        self.norm = normalisation or Normalisation()
        params = params if params is not None else init_parameters(seed)
        self._offsets = {}
        n = 0
        for name, shape in param_shapes():
            size = int(np.prod(shape))
            self._offsets[name] = (n, shape)
            n += size
        self.n_params = n
        f32 = dict(device=self.device, dtype=torch.float32)
        flat = np.concatenate([np.asarray(params[name], np.float32).ravel()
                               for name, _ in param_shapes()])
        self.master = torch.as_tensor(flat, **f32).contiguous()
        self.weights = torch.empty(n, device=self.device, dtype=torch.int16)
        self.grad = torch.zeros(n, **f32)
        self.m = torch.zeros(n, **f32)
        self.v = torch.zeros(n, **f32)
        self.steps = 0
        L.check(self._lib.nfgs_fno_to_bf16(L.ptr(self.master), L.ptr(self.weights), n,
                                           self._stream), "bf16 weights")
        self._bufs = {}

    # -- parameters ---------------------------------------------------------
    @property
    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _w(self, name):
        off, _ = self._offsets[name]
        return L.ptr(self.weights) + 2 * off

    def _m(self, name):
        off, _ = self._offsets[name]
        return L.ptr(self.master) + 4 * off

    def _g(self, name):
        off, _ = self._offsets[name]
        return L.ptr(self.grad) + 4 * off

    def parameters(self, which: str = "master") -> dict:
        """Host copies of the parameters: 'master' (FP32), 'weights' (the
        bfloat16 values the kernels use, as FP32) or 'grad'."""
        if which == "weights":
            raw = self.weights.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
            flat = raw.view(np.float32)
        else:
            flat = (self.master if which == "master" else self.grad).cpu().numpy()
        return {name: flat[off:off + int(np.prod(shape))].reshape(shape).copy()
                for name, (off, shape) in self._offsets.items()}

    def save(self, path):
        np.savez(path, norm=self.norm.as_array(), steps=self.steps,
                 **{k.replace(".", "__"): v for k, v in self.parameters().items()})

    @classmethod
    def load(cls, path, device="cuda"):
        with np.load(path, allow_pickle=False) as z:
            params = {name: z[name.replace(".", "__")] for name, _ in param_shapes()}
            norm = Normalisation(*[float(x) for x in z["norm"]])
        return cls(device=device, params=params, normalisation=norm)

    def _buffers(self, grid, train):
        key = (tuple(grid), train)
        if key not in self._bufs:
            if train:  # a training set of one grid keeps one set of buffers
                self._bufs = {k: b for k, b in self._bufs.items() if not k[1]}
            self._bufs[key] = _Buffers(tuple(grid), self.device, train)
        return self._bufs[key]

    # -- passes ---------------------------------------------------------------
    def features(self, b, dtype_code, B, jp, rho, sb):
        """The network's inputs from a state's B and J_par (solver dtype)."""
        nz, ny, nx = b.grid
        n = self.norm
        L.check(self._lib.nfgs_fno_features(dtype_code, L.ptr(B), L.ptr(jp), nx, ny, nz,
                                            1.0 / (sb * n.b), 1.0 / (sb * n.j), float(rho),
                                            L.ptr(b.feat), self._stream), "learned-start features")

    def forward(self, b):
        """Network output into b.out from b.feat; with training buffers every
        activation is kept for backward."""
        lib, s = self._lib, self._stream
        nz, ny, nx = b.grid
        bzb, byb, bxb = b.bases
        chk = L.check
        chk(lib.nfgs_fno_pointwise(L.ptr(b.feat), FEATURES, self._w("lift.w"), self._m("lift.b"),
                                   None, None, WIDTH, L.NFGS_FNO_IDENTITY, nx, ny, nz,
                                   L.ptr(b.v[0]), s), "lift")
        for l in range(LAYERS):
            vin = b.v[l] if b.train else b.v[l & 1]
            vout = b.v[l + 1] if b.train else b.v[(l + 1) & 1]
            uh = b.uh[l] if b.train else b.uh[0]
            spec = b.spec[l] if b.train else b.spec[0]
            self._spectral_forward(b, l, vin, uh, spec)
            act = L.NFGS_FNO_GELU if l < LAYERS - 1 else L.NFGS_FNO_IDENTITY
            chk(lib.nfgs_fno_pointwise(L.ptr(vin), WIDTH, self._w(f"layer{l}.w"),
                                       self._m(f"layer{l}.b"), L.ptr(spec), L.ptr(bxb), WIDTH, act,
                                       nx, ny, nz, L.ptr(vout), s), f"layer {l}")
        v4 = b.v[LAYERS] if b.train else b.v[LAYERS & 1]
        chk(lib.nfgs_fno_pointwise(L.ptr(v4), WIDTH, self._w("head1.w"), self._m("head1.b"), None,
                                   None, HIDDEN, L.NFGS_FNO_GELU, nx, ny, nz, L.ptr(b.h), s), "head1")
        chk(lib.nfgs_fno_pointwise(L.ptr(b.h), HIDDEN, self._w("head2.w"), self._m("head2.b"), None,
                                   None, OUTPUTS, L.NFGS_FNO_IDENTITY, nx, ny, nz, L.ptr(b.out), s),
            "head2")
        return b.out

    def _spectral_forward(self, b, l, vin, uh, spec):
        lib, s = self._lib, self._stream
        nz, ny, nx = b.grid
        bzb, byb, bxb = b.bases
        P, E = L.NFGS_FNO_PROJECT, L.NFGS_FNO_EXPAND
        chk = L.check
        chk(lib.nfgs_fno_axis(L.ptr(vin), L.ptr(b.ux), L.ptr(bxb), WIDTH * nz * ny, nx, 1, P, s),
            "P_x")
        chk(lib.nfgs_fno_axis(L.ptr(b.ux), L.ptr(b.uy), L.ptr(byb), WIDTH * nz, ny, MODES, P, s),
            "P_y")
        chk(lib.nfgs_fno_axis(L.ptr(b.uy), L.ptr(uh), L.ptr(bzb), WIDTH, nz, MODES * MODES, P, s),
            "P_z")
        chk(lib.nfgs_fno_mix(self._w(f"layer{l}.r"), L.ptr(uh), L.ptr(b.w), NMODES, 0, s), "mix")
        chk(lib.nfgs_fno_axis(L.ptr(b.w), L.ptr(b.sz), L.ptr(bzb), WIDTH, nz, MODES * MODES, E, s),
            "E_z")
        chk(lib.nfgs_fno_axis(L.ptr(b.sz), L.ptr(spec), L.ptr(byb), WIDTH * nz, ny, MODES, E, s),
            "E_y")

    def loss_and_backward(self, b, target):
        """Loss of b.out against target (2, N) and the gradients of every
        parameter, added to self.grad.  Returns the loss as a device double."""
        lib, s = self._lib, self._stream
        nz, ny, nx = b.grid
        N = b.N
        bzb, byb, bxb = b.bases
        P, E = L.NFGS_FNO_PROJECT, L.NFGS_FNO_EXPAND
        chk = L.check
        ws = L.ptr(b.os_ws)
        chk(lib.nfgs_fno_loss(L.ptr(b.out), L.ptr(target), OUTPUTS * N, L.ptr(b.gout),
                              L.ptr(b.loss), L.ptr(b.loss_ws), s), "loss")
        # head
        chk(lib.nfgs_fno_pointwise_backward(L.ptr(b.h), HIDDEN, self._w("head2.w"),
                                            self._m("head2.b"), None, None, OUTPUTS,
                                            L.NFGS_FNO_IDENTITY, L.ptr(b.gout), None, nx, ny, nz,
                                            None, L.ptr(b.gh), s), "head2 backward")
        chk(lib.nfgs_fno_outer_sum(L.ptr(b.gout), OUTPUTS, L.ptr(b.h), HIDDEN, N,
                                   self._g("head2.w"), self._g("head2.b"), 1, ws, s), "head2 grad")
        v4 = b.v[LAYERS]
        chk(lib.nfgs_fno_pointwise_backward(L.ptr(v4), WIDTH, self._w("head1.w"),
                                            self._m("head1.b"), None, None, HIDDEN,
                                            L.NFGS_FNO_GELU, L.ptr(b.gh), None, nx, ny, nz,
                                            L.ptr(b.gz1), L.ptr(b.P[0]), s), "head1 backward")
        chk(lib.nfgs_fno_outer_sum(L.ptr(b.gz1), HIDDEN, L.ptr(v4), WIDTH, N, self._g("head1.w"),
                                   self._g("head1.b"), 1, ws, s), "head1 grad")
        g_full, g_spec = b.P[0], None
        for l in reversed(range(LAYERS)):
            vin = b.v[l]
            gelu_layer = l < LAYERS - 1
            gz = b.gz if gelu_layer else None
            p_out = b.P[(LAYERS - l) & 1]
            chk(lib.nfgs_fno_pointwise_backward(
                L.ptr(vin), WIDTH, self._w(f"layer{l}.w"), self._m(f"layer{l}.b"),
                L.ptr(b.spec[l]) if gelu_layer else None, L.ptr(bxb), WIDTH,
                L.NFGS_FNO_GELU if gelu_layer else L.NFGS_FNO_IDENTITY, L.ptr(g_full),
                L.ptr(g_spec) if g_spec is not None else None, nx, ny, nz,
                L.ptr(gz) if gz is not None else None, L.ptr(p_out), s), f"layer {l} backward")
            gz_t = gz if gz is not None else g_full
            chk(lib.nfgs_fno_outer_sum(L.ptr(gz_t), WIDTH, L.ptr(vin), WIDTH, N,
                                       self._g(f"layer{l}.w"), self._g(f"layer{l}.b"), 1, ws, s),
                f"layer {l} grad")
            gs_out = b.gspec[l & 1]
            chk(lib.nfgs_fno_axis(L.ptr(gz_t), L.ptr(b.ux), L.ptr(bxb), WIDTH * nz * ny, nx, 1, P,
                                  s), "P_x^T")
            chk(lib.nfgs_fno_axis(L.ptr(b.ux), L.ptr(b.uy), L.ptr(byb), WIDTH * nz, ny, MODES, P,
                                  s), "P_y^T")
            chk(lib.nfgs_fno_axis(L.ptr(b.uy), L.ptr(b.gw), L.ptr(bzb), WIDTH, nz, MODES * MODES,
                                  P, s), "P_z^T")
            chk(lib.nfgs_fno_mix_grad(L.ptr(b.gw), L.ptr(b.uh[l]), self._g(f"layer{l}.r"), NMODES,
                                      s), "mix grad")
            chk(lib.nfgs_fno_mix(self._w(f"layer{l}.r"), L.ptr(b.gw), L.ptr(b.guh), NMODES, 1, s),
                "mix^T")
            chk(lib.nfgs_fno_axis(L.ptr(b.guh), L.ptr(b.sz), L.ptr(bzb), WIDTH, nz, MODES * MODES,
                                  E, s), "E_z^T")
            chk(lib.nfgs_fno_axis(L.ptr(b.sz), L.ptr(gs_out), L.ptr(byb), WIDTH * nz, ny, MODES, E,
                                  s), "E_y^T")
            g_full, g_spec = p_out, gs_out
        chk(lib.nfgs_fno_pointwise_backward(None, FEATURES, self._w("lift.w"), self._m("lift.b"),
                                            None, L.ptr(bxb), WIDTH, L.NFGS_FNO_IDENTITY,
                                            L.ptr(g_full), L.ptr(g_spec), nx, ny, nz,
                                            L.ptr(b.gz), None, s), "lift backward")
        chk(lib.nfgs_fno_outer_sum(L.ptr(b.gz), WIDTH, L.ptr(b.feat), FEATURES, N,
                                   self._g("lift.w"), self._g("lift.b"), 1, ws, s), "lift grad")
        return b.loss

    def adam_step(self, lr, batch: int = 1, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam step over the gradients accumulated from `batch` samples
        (clears them)."""
        self.steps += 1
        L.check(self._lib.nfgs_fno_adam(L.ptr(self.master), L.ptr(self.weights), L.ptr(self.m),
                                        L.ptr(self.v), L.ptr(self.grad), self.n_params, float(lr),
                                        beta1, beta2, eps, self.steps, 1.0 / batch, self._stream),
                "adam")

    # -- the solver-facing start ---------------------------------------------
    def predict_current(self, solver, B, jp):
        """jp (the gs_cfcf state's J_par, solver dtype, 3 x N) += the network's
        correction; B is that state's field."""
        grid = (solver.Nz, solver.Ny, solver.Nx)
        b = self._buffers(grid, False)
        code = L.NFGS_F64 if solver.dtype == torch.float64 else L.NFGS_F32
        sb = _bz_scale(solver)
        self.features(b, code, B, jp, solver.bottom_divergence_robin_weight, sb)
        out = self.forward(b)
        L.check(self._lib.nfgs_fno_apply(code, L.ptr(jp), L.ptr(out), OUTPUTS * b.N,
                                         sb * self.norm.out, L.ptr(jp), self._stream),
                "learned start")


# ---------------------------------------------------------------------------
# Training pairs from the solver, training and evaluation
# ---------------------------------------------------------------------------
@dataclass
class Pair:
    """One (boundary, rho) -> converged-current pair on the device (FP32):
    the gs_cfcf state's B and J_par and the selected state's J_par."""
    b_gs: torch.Tensor
    jp_gs: torch.Tensor
    jp_conv: torch.Tensor
    rho: float
    sb: float
    iterations: int
    seconds: float


def solve_pair(solver, rho, iterate_kw) -> Pair:
    """Plain solve of a solver whose boundary is set; returns its pair."""
    t0 = time.perf_counter()
    B, jp = solver.gs_cfcf_start()
    f = lambda t: t.to(torch.float32).contiguous().clone()
    b_gs, jp_gs = f(B), f(jp)
    h = solver.iterate(**iterate_kw)
    jp_c = f(solver.parallel_current())
    torch.cuda.synchronize(solver.device)
    return Pair(b_gs, jp_gs, jp_c, float(rho), _bz_scale(solver), len(h["sigma"]),
                time.perf_counter() - t0)


def fit_normalisation(pairs) -> Normalisation:
    """Channel scales from the training pairs (RMS over the set, in units of
    each pair's s_B)."""
    acc = np.zeros(3)
    for p in pairs:
        vals = [p.b_gs.double(), p.jp_gs.double(), (p.jp_conv - p.jp_gs).double()]
        acc += np.array([float(torch.mean(v * v).item()) / p.sb ** 2 for v in vals])
    rms = np.sqrt(acc / max(len(pairs), 1))
    return Normalisation(*[float(max(r, 1e-12)) for r in rms])


def pair_target(net: SpectralWarmStart, p: Pair, out: torch.Tensor):
    L.check(net._lib.nfgs_fno_targets(L.NFGS_F32, L.ptr(p.jp_gs), L.ptr(p.jp_conv),
                                      p.jp_gs.numel(), 1.0 / (p.sb * net.norm.out), L.ptr(out),
                                      net._stream), "targets")


def _pair_forward(net, b, p):
    net.features(b, L.NFGS_F32, p.b_gs, p.jp_gs, p.rho, p.sb)
    net.forward(b)


def train(net: SpectralWarmStart, pairs, epochs: int, batch: int = 4, lr: float = 1e-3,
          lr_final: float = 1e-5, seed: int = 0, time_limit: Optional[float] = None, log=None):
    """Adam over the pairs with a cosine learning-rate schedule.  Returns the
    mean loss per epoch (read once per epoch)."""
    grid = tuple(pairs[0].b_gs.shape[1:])
    b = net._buffers(grid, True)
    target = torch.empty(OUTPUTS, b.N, device=net.device, dtype=torch.float32)
    rng = np.random.default_rng(seed)
    steps_total = max(1, epochs * math.ceil(len(pairs) / batch))
    history = []
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        losses = torch.zeros(1, device=net.device, dtype=torch.float64)
        for k0 in range(0, len(order), batch):
            idx = order[k0:k0 + batch]
            for i in idx:
                p = pairs[int(i)]
                _pair_forward(net, b, p)
                pair_target(net, p, target)
                losses += net.loss_and_backward(b, target)
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * step / steps_total))
            net.adam_step(lr_t, batch=len(idx))
            step += 1
        history.append(float(losses.item()) / len(order))
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {history[-1]:.4e} "
                f"({time.perf_counter() - t0:.1f} s)")
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            break
    return history


def evaluate_loss(net: SpectralWarmStart, pairs) -> float:
    """Mean loss of the network over pairs (no parameter change)."""
    grid = tuple(pairs[0].b_gs.shape[1:])
    b = net._buffers(grid, True)
    target = torch.empty(OUTPUTS, b.N, device=net.device, dtype=torch.float32)
    tot = torch.zeros(1, device=net.device, dtype=torch.float64)
    for p in pairs:
        _pair_forward(net, b, p)
        pair_target(net, p, target)
        L.check(net._lib.nfgs_fno_loss(L.ptr(b.out), L.ptr(target), OUTPUTS * b.N, L.ptr(b.gout),
                                       L.ptr(b.loss), L.ptr(b.loss_ws), net._stream), "loss")
        tot += b.loss
    return float(tot.item()) / len(pairs)


# ---------------------------------------------------------------------------
# Solver-generated pairs of the multipole family and the plain-vs-learned
# comparison
# ---------------------------------------------------------------------------
# The production profile of the reference's controlled runs
# (scripts/run_controlled_native.py:25-31) and its paper stopping protocol.
SOLVER_KW = dict(lateral_wall="paper_bc1", pin_potential_boundaries=False, buffer_thickness=0,
                 padded_buffer_zone=True, pad_calculation_domain=True,
                 calculation_domain_mode="mudpack", mean_bz_policy="uniform_open",
                 horizontal_zero_mode_policy="frozen_linear_harmonic",
                 d0_zero_mode_policy="project", source_zero_mode_policy="project",
                 harmonic_residual_policy="diagnose", poloidal_update_mode="direct",
                 vertical_current_operator="staggered_face")
ITERATE_KW = dict(max_iter=500, tol=1e-5, mu=1, patience=40, sigma_improvement_tol=1e-5,
                  stop_mode="sigma_best", verbose=False)


def case(n: int, seed: int, rho: Optional[float] = None):
    """Boundary planes (device FP64) and rho of one case of the family: the
    seeded multipole boundary on n x n observed columns (pole widths 8 to 20
    cells from n = 128 up, scaled by n / 128 below), rho ~ U[0, 1] from the
    same seed unless given."""
    from .boundary import multipole_boundary
    f = min(1.0, n / 128.0)
    bx, by, bz, _ = multipole_boundary(n, seed=seed, sigma=(8.0 * f, 20.0 * f))
    if rho is None:
        rho = float(np.random.default_rng(10_000 + seed).uniform(0.0, 1.0))
    return tuple(torch.as_tensor(p, device="cuda") for p in (bx, by, bz)), float(rho)


def make_solver(n: int, planes, rho: float, nz: Optional[int] = None,
                dtype=torch.float32):
    from .solver import NFGSSolverGPU
    s = NFGSSolverGPU(n, n, n if nz is None else nz, 1.0, 1.0, 1.0, device="cuda", dtype=dtype,
                      bottom_divergence_robin_weight=rho, **SOLVER_KW)
    s.set_bottom_vector_boundary(*(p.to(dtype) for p in planes), zero_net_current=False,
                                 smooth_passes=0, ijzbcor=3)
    return s


def generate_pairs(count: int, n: int = 128, seed0: int = 0, nz: Optional[int] = None,
                   log=None):
    """`count` solver-generated pairs of the multipole family (seeds seed0,
    seed0 + 1, ...), each a plain FP32 solve to the sigma_best stop."""
    pairs = []
    t0 = time.perf_counter()
    for i in range(count):
        planes, rho = case(n, seed0 + i)
        s = make_solver(n, planes, rho, nz)
        pairs.append(solve_pair(s, rho, ITERATE_KW))
        del s
        if log and (i + 1) % max(1, count // 10) == 0:
            log(f"pairs {i + 1}/{count} ({time.perf_counter() - t0:.1f} s)")
    return pairs


def compare_starts(net: SpectralWarmStart, n: int, seeds, nz: Optional[int] = None):
    """Plain gs_cfcf solve and learned-start solve of each held-out case:
    iterations, wall time (set-up, start and iteration, synchronised), best
    sigma_J and the relative L2 distance of the selected B between them."""
    rows = []
    for seed in seeds:
        planes, rho = case(n, seed)
        res = {}
        for mode in ("gs_cfcf", "learned"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s = make_solver(n, planes, rho, nz)
            kw = dict(ITERATE_KW, init_mode=mode)
            if mode == "learned":
                kw["learned_start"] = net
            h = s.iterate(**kw)
            torch.cuda.synchronize()
            res[mode] = (s, h, time.perf_counter() - t0)
        (sp, hp, tp), (sl, hl, tl) = res["gs_cfcf"], res["learned"]
        db = float(torch.linalg.norm((sl._B - sp._B).double()) /
                   torch.linalg.norm(sp._B.double()))
        rows.append(dict(seed=int(seed), rho=rho, plain_iterations=len(hp["sigma"]),
                         learned_iterations=len(hl["sigma"]), plain_seconds=tp,
                         learned_seconds=tl, plain_best_sigma=hp["best_sigma"],
                         learned_best_sigma=hl["best_sigma"],
                         learned_used=hl["init_diagnostics"]["learned_start_used"],
                         sigma_gs_cfcf=hl["init_diagnostics"]["sigma_gs_cfcf"],
                         sigma_learned=hl["init_diagnostics"]["sigma_learned"],
                         B_relative_to_plain=db))
        del res
    return rows


def summarize(rows) -> dict:
    it_p = np.array([r["plain_iterations"] for r in rows], float)
    it_l = np.array([r["learned_iterations"] for r in rows], float)
    t_p = np.array([r["plain_seconds"] for r in rows])
    t_l = np.array([r["learned_seconds"] for r in rows])
    sg = np.array([r["learned_best_sigma"] / r["plain_best_sigma"] - 1.0 for r in rows])
    return dict(cases=len(rows), used=int(sum(r["learned_used"] for r in rows)),
                plain_iterations_mean=float(it_p.mean()),
                learned_iterations_mean=float(it_l.mean()),
                wall_speedup_total=float(t_p.sum() / t_l.sum()),
                wall_speedup_median=float(np.median(t_p / t_l)),
                best_sigma_change_median=float(np.median(sg)),
                best_sigma_change_max=float(sg.max()),
                B_relative_to_plain_median=float(np.median([r["B_relative_to_plain"]
                                                            for r in rows])))


def main(argv=None):
    """python -m nfgs_amd.learned --pairs P --n N --epochs E [--out model.npz]:
    generate P solver pairs, train, and compare plain and learned starts on
    held-out seeds (one JSON line)."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--pairs", type=int, default=200)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--holdout", type=int, default=16)
    ap.add_argument("--seed0", type=int, default=0)
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    t0 = time.perf_counter()
    pairs = generate_pairs(a.pairs, a.n, a.seed0, log=log)
    t_gen = time.perf_counter() - t0
    net = SpectralWarmStart(seed=a.seed0)
    net.norm = fit_normalisation(pairs)
    t1 = time.perf_counter()
    hist = train(net, pairs, a.epochs, batch=a.batch, lr=a.lr, time_limit=a.time_limit, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    if a.out:
        net.save(a.out)
    del pairs
    rows = compare_starts(net, a.n, range(a.seed0 + 1_000_000, a.seed0 + 1_000_000 + a.holdout))
    print(json.dumps(dict(pairs=a.pairs, n=a.n, epochs=len(hist), loss=hist,
                          generate_seconds=t_gen, train_seconds=t_train,
                          summary=summarize(rows), cases=rows)), flush=True)


if __name__ == "__main__":
    main()
