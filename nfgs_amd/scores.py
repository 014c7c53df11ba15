"""Common scoring of output fields on the GPU (FP64).

Same names, arguments and results as the reference's scoring helpers:
`curl`, `divergence` and `sigma_j` of nfgs/diagnostics/common_fd2.py:41-83,
`roi_cwsin`, `vector_correlation` and `compute_reference_metrics` of
nfgs/_backend/util/metrics.py:55-176, and `compute_metrics` of
nfgs/_backend/scripts/recompute_td_referee_diagnostics.py:344-400 (the
controlled evaluation's metrics).  Fields are dicts of (nz, ny, nx)
arrays (NumPy or torch); the arithmetic runs in the HIP kernels of
libnfgs_hip.so (nfgs_fd2_curl_div, nfgs_field_scores) in FP64, as the
reference evaluates its common diagnostics.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L

_B = ("Bx", "By", "Bz")
_J = ("Jx", "Jy", "Jz")


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("nfgs_amd.scores needs a ROCm GPU: the HIP library has no CPU path")
    return torch.device("cuda", torch.cuda.current_device())


def _vec(parts):
    """(3, nz, ny, nx) FP64 device tensor from three arrays."""
    dev = _device()
    ts = [torch.as_tensor(np.asarray(p) if not torch.is_tensor(p) else p) for p in parts]
    shape = tuple(ts[0].shape)
    if len(shape) != 3 or any(tuple(t.shape) != shape for t in ts):
        raise ValueError(f"components must be (nz, ny, nx) arrays of one shape; got "
                         f"{[tuple(t.shape) for t in ts]}")
    return torch.stack([t.to(device=dev, dtype=torch.float64) for t in ts]).contiguous()


def _vol(a):
    """(nz, ny, nx) FP64 device tensor from one array."""
    t = torch.as_tensor(np.asarray(a) if not torch.is_tensor(a) else a)
    if t.dim() != 3:
        raise ValueError(f"expected an (nz, ny, nx) array; got {tuple(t.shape)}")
    return t.to(device=_device(), dtype=torch.float64).contiguous()


def _spacing_xyz(spacing):
    """An isotropic scalar or physical spacings ordered x, y, z (common_fd2.py:11-20)."""
    values = (float(spacing),) * 3 if np.isscalar(spacing) else tuple(float(v) for v in spacing)
    if len(values) != 3 or any(not np.isfinite(v) or v <= 0 for v in values):
        raise ValueError("spacing must be positive: a scalar or (dx, dy, dz)")
    return values


def _box(slices, shape):
    out = []
    for sl, n in zip(slices, shape):
        start, stop, step = sl.indices(n)
        if step != 1 or stop <= start:
            raise ValueError("scoring slices must be non-empty and unstrided")
        out += [start, stop]
    return out


def _curl_div(B, spacing, want_j=True, want_div=False):
    lib = L.load()
    _, nz, ny, nx = B.shape
    dx, dy, dz = _spacing_xyz(spacing)
    j = torch.empty_like(B) if want_j else None
    d = torch.empty_like(B[0]) if want_div else None
    L.check(lib.nfgs_fd2_curl_div(L.ptr(B), nz, ny, nx, dx, dy, dz,
                                  L.ptr(j) if j is not None else None,
                                  L.ptr(d) if d is not None else None,
                                  torch.cuda.current_stream().cuda_stream), "fd2 curl/div")
    return j, d


def _sums(bref, b, j, slices, b2min, div=None):
    lib = L.load()
    _, nz, ny, nx = b.shape
    box = _box(slices, (nz, ny, nx))
    ws = torch.empty(int(lib.nfgs_scores_workspace_bytes()) // 8, device=b.device,
                     dtype=torch.float64)
    out = torch.zeros(L.SC_COUNT, device=b.device, dtype=torch.float64)
    opt = [L.ptr(t) if t is not None else None for t in (bref, j, div)]
    L.check(lib.nfgs_field_scores(opt[0], L.ptr(b), opt[1], opt[2], nz, ny, nx, *box,
                                  float(b2min), L.ptr(ws), L.ptr(out),
                                  torch.cuda.current_stream().cuda_stream), "field scores")
    s = out.cpu().numpy()
    count = float(np.prod([box[1] - box[0], box[3] - box[2], box[5] - box[4]]))
    return s, count


def curl(fields, spacing):
    """Second-order curl of the B components (common_fd2.py:41-49); NumPy out."""
    j, _ = _curl_div(_vec([fields[k] for k in _B]), spacing)
    jh = j.cpu().numpy()
    return {k: jh[i] for i, k in enumerate(_J)}


def divergence(fields, spacing):
    """Second-order divergence (common_fd2.py:52-61); NumPy out."""
    _, d = _curl_div(_vec([fields[k] for k in _B]), spacing, want_j=False, want_div=True)
    return d.cpu().numpy()


def curl_and_divergence(fields, spacing):
    """curl and divergence of one field in a single pass (NumPy out)."""
    j, d = _curl_div(_vec([fields[k] for k in _B]), spacing, want_j=True, want_div=True)
    jh = j.cpu().numpy()
    return {k: jh[i] for i, k in enumerate(_J)}, d.cpu().numpy()


def sigma_j(fields, spacing, slices) -> float:
    """sigma_J of the common curl on a box (common_fd2.py:64-83)."""
    B = _vec([fields[k] for k in _B])
    j, _ = _curl_div(B, spacing)
    s, _ = _sums(None, B, j, slices, 1.0e-10)
    return float(s[L.SC_JXB] / max(float(s[L.SC_JN]), 1.0e-30))


def roi_cwsin(B_vec, J_vec, *, b2min: float = 1e-10) -> float:
    """Current-weighted sine on pre-sliced (3, nz, ny, nx) volumes (metrics.py:66-92)."""
    B = _vec(list(B_vec))
    J = _vec(list(J_vec))
    if B.shape != J.shape:
        raise ValueError(f"B_vec/J_vec must both have shape (3, Nz, Ny, Nx); got "
                         f"{tuple(B.shape)}, {tuple(J.shape)}")
    s, _ = _sums(None, B, J, (slice(None),) * 3, b2min)
    return float(s[L.SC_JXB] / max(float(s[L.SC_JN]), 1e-30))


def vector_correlation(pred, ref) -> float:
    """Global vector correlation of two (3, nz, ny, nx) volumes (metrics.py:55-59)."""
    b = _vec(list(pred))
    r = _vec(list(ref))
    s, _ = _sums(r, b, None, (slice(None),) * 3, 1e-10)
    return float(s[L.SC_DOT]) / (float(np.sqrt(s[L.SC_SOL2] * s[L.SC_REF2])) + 1e-12)


def compute_reference_metrics(ref_fields, fields, metric_slices):
    """Reference-field agreement and sigma_J on a z-y-x box (metrics.py:128-176)."""
    R = _vec([ref_fields[k] for k in _B])
    S = _vec([fields[k] for k in _B])
    J = _vec([fields[k] for k in _J])
    if R.shape != S.shape or J.shape != S.shape:
        raise ValueError("reference and solution fields must have one shape")
    s, count = _sums(R, S, J, metric_slices, 1e-10)
    if not np.all(np.isfinite(s[:L.SC_JXB])):
        raise RuntimeError("Non-finite values in B fields for metrics.")
    if not np.all(np.isfinite(s[L.SC_JXB:])):
        raise RuntimeError("Non-finite values in J for sigma_J.")
    c_vec = float(s[L.SC_DOT]) / (float(np.sqrt(s[L.SC_SOL2] * s[L.SC_REF2])) + 1e-12)
    e_n = float(s[L.SC_DIFF] / (s[L.SC_REFN] + 1e-12))
    e_m = float(s[L.SC_DIFF_REL] / count)
    return {"C_vec": c_vec, "C_CS": float(s[L.SC_COS] / count), "En": e_n, "Em": e_m,
            "Eprime_n": 1.0 - e_n, "Eprime_m": 1.0 - e_m,
            "epsilon": float(s[L.SC_SOL2] / (s[L.SC_REF2] + 1e-12)),
            "sigma_J": float(s[L.SC_JXB] / max(float(s[L.SC_JN]), 1e-30))}


def compute_metrics(reference, candidate, current, divergence, slices, spacing) -> dict:
    """Agreement metrics, sigma_J and the normalised divergence of a candidate
    on a z-y-x box (recompute_td_referee_diagnostics.py:344-400); `current`
    holds Jx, Jy, Jz, `divergence` is one volume and `spacing` the effective
    grid step of the divergence normalisation."""
    R = _vec([reference[k] for k in _B])
    S = _vec([candidate[k] for k in _B])
    J = _vec([current[k] for k in _J])
    D = _vol(divergence)
    if R.shape != S.shape or J.shape != S.shape or tuple(D.shape) != tuple(S.shape[1:]):
        raise ValueError("reference, candidate, current and divergence must have one shape")
    s, count = _sums(R, S, J, slices, 1.0e-10, div=D)
    n = max(count, 1.0)
    div_mean = float(s[L.SC_DIVABS]) / n
    bn_mean = float(s[L.SC_BN]) / n
    return {
        "C_vec": float(s[L.SC_DOT]) / (float(np.sqrt(s[L.SC_REF2] * s[L.SC_SOL2])) + 1.0e-12),
        "C_CS": float(s[L.SC_COS]) / n,
        "Eprime_n": 1.0 - float(s[L.SC_DIFF]) / (float(s[L.SC_REFN]) + 1.0e-12),
        "Eprime_m": 1.0 - float(s[L.SC_DIFF_REL]) / n,
        "epsilon": float(s[L.SC_SOL2]) / (float(s[L.SC_REF2]) + 1.0e-12),
        "sigma_J": float(s[L.SC_JXB]) / max(float(s[L.SC_JN]), 1.0e-30),
        "divB_normalized": div_mean / (6.0 * bn_mean / float(spacing) + 1.0e-12),
    }
