"""Compact digests of large arrays, for parity fixtures at full grid sizes.

A 258x258x257 FP64 volume is 137 MB, too large to commit, so full-size
oracle runs are stored as digests and the GPU arrays are digested the same
way at test time.  Per component (arrays are viewed as (C, nz, ny, nx)):

  sq_z, sq_y, sq_x   sums of squares over the other two axes
  sum_z              signed sums over each z-plane
  sketch             16 projections on separable random +-1 patterns
                     s(z) s(y) s(x); E[(sketch of d)^2] = ||d||^2, so the
                     sketch of a difference estimates its L2 norm
  sample             values at 1024 fixed pseudo-random points

`compare` turns two digests into relative errors, each normalised by the
reference component's L2 norm; a wrong plane, row or column moves a
marginal, a diffuse error moves the sketches and the samples.
"""
from __future__ import annotations

import numpy as np

NSKETCH = 16
NSAMPLE = 1024
_SEED = 20261204


def _as4(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 2:
        return a[None, None]
    if a.ndim == 3:
        return a[None]
    if a.ndim == 4:
        return a
    raise ValueError(f"cannot digest an array of shape {a.shape}")


def _patterns(nz, ny, nx):
    rng = np.random.default_rng(_SEED + 7 * nz + 11 * ny + 13 * nx)
    sz = rng.choice([-1.0, 1.0], size=(NSKETCH, nz))
    sy = rng.choice([-1.0, 1.0], size=(NSKETCH, ny))
    sx = rng.choice([-1.0, 1.0], size=(NSKETCH, nx))
    idx = rng.integers(0, nz * ny * nx, size=NSAMPLE)
    return sz, sy, sx, idx


def digest(a) -> dict:
    """Digest of one array (a plane, a volume or a stack of components)."""
    a = _as4(a)
    c, nz, ny, nx = a.shape
    sz, sy, sx, idx = _patterns(nz, ny, nx)
    sq = a * a
    return {
        "shape": np.array(a.shape),
        "sq_z": sq.sum(axis=(2, 3)),
        "sq_y": sq.sum(axis=(1, 3)),
        "sq_x": sq.sum(axis=(1, 2)),
        "sum_z": a.sum(axis=(2, 3)),
        "sketch": np.einsum("czyx,kz,ky,kx->ck", a, sz, sy, sx, optimize=True),
        "sample": a.reshape(c, -1)[:, idx],
    }


def compare(got: dict, want: dict) -> dict:
    """Relative errors of `got` against `want`, per component (arrays of C
    values): sketch (estimated relative L2), sample (relative L2 over the
    sampled points), plane/row/column norms and plane sums."""
    if tuple(got["shape"]) != tuple(want["shape"]):
        raise AssertionError(f"digest shapes differ: {got['shape']} vs {want['shape']}")
    c, nz, ny, nx = (int(v) for v in want["shape"])
    norm = np.sqrt(want["sq_z"].sum(axis=1))
    norm = np.where(norm > 0, norm, 1.0)
    out = {
        "sketch": np.sqrt(np.mean((got["sketch"] - want["sketch"]) ** 2, axis=1)) / norm,
        "sample": np.linalg.norm(got["sample"] - want["sample"], axis=1) /
        np.maximum(np.linalg.norm(want["sample"], axis=1), 1e-300),
    }
    for key in ("sq_z", "sq_y", "sq_x"):
        out[key] = np.max(np.abs(np.sqrt(got[key]) - np.sqrt(want[key])), axis=1) / norm
    out["sum_z"] = np.max(np.abs(got["sum_z"] - want["sum_z"]), axis=1) / \
        (np.sqrt(ny * nx) * norm)
    return out


def worst(got: dict, want: dict) -> float:
    """Largest relative error of compare() over every measure and component."""
    return float(max(np.max(v) for v in compare(got, want).values()))


def pack(prefix: str, d: dict) -> dict:
    """Flatten a digest into NPZ entries prefix/key."""
    return {f"{prefix}/{k}": v for k, v in d.items()}


def unpack(npz: dict, prefix: str) -> dict:
    p = prefix + "/"
    return {k[len(p):]: v for k, v in npz.items() if k.startswith(p)}
