"""Learned warm start for the NFGS iterate (an extension; the reference starts
only from gs_cfcf, nfgs_solver.py:6613-6645).

A small cosine-basis spectral neural operator maps what the solver knows
before its first step (the gs_cfcf scalars Phi_c and Az, the bottom planes Bx,
By, Bz, Jz on the calculation grid and the Robin weight rho) to a correction
of (Phi_c, Az) towards the converged state.  ``iterate(init_mode="learned",
learned_start=net)`` evaluates sigma_J of both starts and keeps the learned
one only when its sigma_J is lower; otherwise the solve is the plain gs_cfcf
solve.  The iteration itself is unchanged, so the converged result is judged
by the same monitors and stop rules.

Network: lift 8 -> 32, four spectral layers (truncated orthonormal DCT-II of
16 modes per axis, a 32 x 32 mix per mode, a pointwise 32 x 32 path, GELU
between layers), head 32 -> 64 -> 2.  Every pass runs in the library's
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

WIDTH, MODES, FEATURES, HIDDEN, OUTPUTS, LAYERS = 32, 16, 8, 64, 2, 4
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
    """Fixed channel scales, in units of the boundary's RMS Bz (s_B): the
    gs_cfcf scalars are divided by s_B * phi / az, Jz by s_B * jz, the
    targets by s_B * out_phi / out_az."""
    phi: float = 1.0
    az: float = 1.0
    jz: float = 1.0
    out_phi: float = 1.0
    out_az: float = 1.0

    def as_array(self):
        return np.array([self.phi, self.az, self.jz, self.out_phi, self.out_az])


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
            ws = max(int(L.load().nfgs_fno_outer_sum_workspace_bytes(a, b))
                     for a, b in ((WIDTH, FEATURES), (WIDTH, WIDTH), (HIDDEN, WIDTH),
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
    def features(self, b, dtype_code, phi, az, bx, by, bz, jz, rho, sb):
        nz, ny, nx = b.grid
        n = self.norm
        L.check(self._lib.nfgs_fno_features(
            dtype_code, L.ptr(phi), L.ptr(az), L.ptr(bx), L.ptr(by), L.ptr(bz), L.ptr(jz),
            nx, ny, nz, 1.0 / (sb * n.phi), 1.0 / (sb * n.az), 1.0 / sb, 1.0 / (sb * n.jz),
            float(rho), L.ptr(b.feat), self._stream), "learned-start features")

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
        chk(lib.nfgs_fno_loss(L.ptr(b.out), L.ptr(target), nx, ny, nz, L.ptr(b.gout),
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
    def predict_start(self, solver, gs_phi, gs_az, phi_out, az_out):
        """(phi_out, az_out) = the gs_cfcf state (gs_phi, gs_az) plus the
        network's correction, on the solver's grid and dtype."""
        grid = (solver.Nz, solver.Ny, solver.Nx)
        b = self._buffers(grid, False)
        code = L.NFGS_F64 if solver.dtype == torch.float64 else L.NFGS_F32
        sb = _bz_scale(solver)
        self.features(b, code, gs_phi, gs_az, solver.Bx_obs, solver.By_obs, solver.Bz_obs,
                      solver.Jz_obs, solver.bottom_divergence_robin_weight, sb)
        out = self.forward(b)
        nz, ny, nx = grid
        L.check(self._lib.nfgs_fno_apply(code, L.ptr(gs_phi), L.ptr(gs_az), L.ptr(out), nx, ny, nz,
                                         sb * self.norm.out_phi, sb * self.norm.out_az,
                                         L.ptr(phi_out), L.ptr(az_out), self._stream),
                "learned start")


# ---------------------------------------------------------------------------
# Training pairs from the solver, training and evaluation
# ---------------------------------------------------------------------------
@dataclass
class Pair:
    """One (boundary, rho) -> converged-state pair on the device (FP32)."""
    gs_phi: torch.Tensor
    gs_az: torch.Tensor
    conv_phi: torch.Tensor
    conv_az: torch.Tensor
    planes: tuple  # Bx, By, Bz, Jz on the calculation grid
    rho: float
    sb: float
    iterations: int
    seconds: float


def solve_pair(solver, rho, iterate_kw) -> Pair:
    """Plain solve of a solver whose boundary is set; returns its pair."""
    t0 = time.perf_counter()
    _, _, phi0, az0, _ = solver.compute_initial_conditions()
    h = solver.iterate(**iterate_kw)
    torch.cuda.synchronize(solver.device)
    f = lambda t: t.to(torch.float32).contiguous()
    return Pair(f(phi0), f(az0), f(solver._Phi_c), f(solver._Az),
                tuple(f(p) for p in (solver.Bx_obs, solver.By_obs, solver.Bz_obs, solver.Jz_obs)),
                float(rho), _bz_scale(solver), len(h["sigma"]), time.perf_counter() - t0)


def fit_normalisation(pairs) -> Normalisation:
    """Channel scales from the training pairs (RMS over the set, in units of
    each pair's s_B)."""
    acc = np.zeros(5)
    for p in pairs:
        d_phi = (p.conv_phi - p.gs_phi).double()
        d_az = (p.conv_az - p.gs_az).double()
        vals = [p.gs_phi.double(), p.gs_az.double(), p.planes[3].double(), d_phi, d_az]
        acc += np.array([float(torch.mean(v * v).item()) / p.sb ** 2 for v in vals])
    rms = np.sqrt(acc / max(len(pairs), 1))
    return Normalisation(*[float(max(r, 1e-12)) for r in rms])


def pair_target(net: SpectralWarmStart, p: Pair, out: torch.Tensor):
    nz, ny, nx = p.gs_phi.shape
    n = net.norm
    L.check(net._lib.nfgs_fno_targets(L.NFGS_F32, L.ptr(p.gs_phi), L.ptr(p.gs_az),
                                      L.ptr(p.conv_phi), L.ptr(p.conv_az), nx, ny, nz,
                                      1.0 / (p.sb * n.out_phi), 1.0 / (p.sb * n.out_az),
                                      L.ptr(out), net._stream), "targets")


def train(net: SpectralWarmStart, pairs, epochs: int, batch: int = 4, lr: float = 1e-3,
          lr_final: float = 1e-5, seed: int = 0, time_limit: Optional[float] = None, log=None):
    """Adam over the pairs with a cosine learning-rate schedule.  Returns the
    mean loss per epoch (read once per epoch)."""
    grid = tuple(pairs[0].gs_phi.shape)
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
                net.features(b, L.NFGS_F32, p.gs_phi, p.gs_az, *p.planes, p.rho, p.sb)
                net.forward(b)
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
    grid = tuple(pairs[0].gs_phi.shape)
    b = net._buffers(grid, True)
    target = torch.empty(OUTPUTS, b.N, device=net.device, dtype=torch.float32)
    tot = torch.zeros(1, device=net.device, dtype=torch.float64)
    lib = net._lib
    for p in pairs:
        net.features(b, L.NFGS_F32, p.gs_phi, p.gs_az, *p.planes, p.rho, p.sb)
        net.forward(b)
        pair_target(net, p, target)
        nz, ny, nx = grid
        L.check(lib.nfgs_fno_loss(L.ptr(b.out), L.ptr(target), nx, ny, nz, L.ptr(b.gout),
                                  L.ptr(b.loss), L.ptr(b.loss_ws), net._stream), "loss")
        tot += b.loss
    return float(tot.item()) / len(pairs)
