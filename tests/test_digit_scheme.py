"""The arithmetic of the FP32 int8 transform, emulated in numpy (no GPU).

paired_transform_i8 (nfgs_amd/csrc/transform.hip) scales each vector by a
power of two to int32 fixed point (|X| < 2^30), splits X and the matrix
entries round(T * 2^shift) into four signed base-256 digits with
(X + 0x808080) ^ 0x808080, multiplies the ten digit pairs of weight
>= 256^-3 exactly in integers and combines the four weight levels in FP64.
These tests pin that scheme against FP64 products of the type-I matrices
(the exact operators of nfgs_solver.py:1019-1081): the digits reconstruct X
exactly, the ten pairs reach FP64-arithmetic accuracy, and six pairs do not.
"""
import numpy as np
import pytest

from oracle import nfgs_oracle as O


def _digits(X):
    D = ((X.astype(np.int64) + 0x00808080) ^ 0x00808080) & 0xFFFFFFFF
    out = []
    for p in range(4):
        b = (D >> (8 * (3 - p))) & 0xFF
        out.append(np.where(b >= 128, b - 256, b))
    return out


def _type1(kind, n):
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    if kind == "dct":
        w = np.where((j == 0) | (j == n - 1), 1.0, 2.0)
        return w * np.cos(np.pi * ((j * k) % (2 * (n - 1))) / (n - 1))
    return 2.0 * np.sin(np.pi * (((j + 1) * (k + 1)) % (2 * (n + 1))) / (n + 1))


def _digit_product(M, X, pairs):
    """M (n x n) times the FP32 columns of X the way the kernel does it."""
    mx = np.abs(X).max(axis=0).astype(np.float64)
    e = np.frexp(np.where(mx > 0, mx, 1.0))[1]  # max|x| < 2^e
    s = 30 - e
    Xi = np.trunc(X.astype(np.float64) * 2.0 ** s).astype(np.int64)
    a = _digits(Xi)
    sh = 30 - np.frexp(np.abs(M).max())[1]
    if np.abs(M).max() * 2.0 ** (sh + 1) <= 2.0 ** 30:
        sh += 1
    c = _digits(np.rint(M * 2.0 ** sh).astype(np.int64))
    acc = np.zeros((M.shape[0], X.shape[1]))
    for i, j in pairs:
        acc += (c[i] @ a[j]).astype(np.float64) * 256.0 ** (6 - i - j)
    return (acc * 2.0 ** -sh * 2.0 ** -s).astype(np.float32)


TEN = [(i, j) for i in range(4) for j in range(4) if i + j <= 3]
SIX = [(i, j) for i in range(3) for j in range(3) if i + j <= 2]


def test_digits_reconstruct_exactly():
    rng = np.random.default_rng(0)
    X = rng.integers(-(2 ** 30), 2 ** 30, size=10000)
    d = _digits(X)
    assert all(np.abs(v).max() <= 128 for v in d)
    rec = sum(d[p] * 256 ** (3 - p) for p in range(4))
    assert np.array_equal(rec, X)


@pytest.mark.parametrize("kind,n", [("dct", 258), ("dct", 130), ("dst", 256)])
@pytest.mark.parametrize("case", ["walk", "gauss", "range"])
def test_ten_pairs_match_fp64_products(kind, n, case):
    rng = np.random.default_rng(n)
    if case == "walk":
        X = np.cumsum(rng.standard_normal((n, 64)), axis=0)
    elif case == "gauss":
        X = rng.standard_normal((n, 64))
    else:
        X = rng.standard_normal((n, 64)) * 10.0 ** rng.uniform(-6, 0, (n, 64))
    X = X.astype(np.float32)
    M = _type1(kind, n)
    ref = M @ X.astype(np.float64)
    fp64_out = ref.astype(np.float32)  # FP64 arithmetic, one FP32 rounding
    e64 = np.linalg.norm(fp64_out - ref) / np.linalg.norm(ref)
    e10 = np.linalg.norm(_digit_product(M, X, TEN) - ref) / np.linalg.norm(ref)
    e6 = np.linalg.norm(_digit_product(M, X, SIX) - ref) / np.linalg.norm(ref)
    assert e10 < 1.1 * e64 + 1e-9, (e10, e64)
    assert e6 > 5 * e64, (e6, e64)  # levels <= 2 alone are not enough
    # and the FFT form the reference uses for its FP32 transforms
    f = O.dct1 if kind == "dct" else O.dst1
    eff = np.linalg.norm(f(X, 0).astype(np.float64) - ref) / np.linalg.norm(ref)
    assert e10 < 1.5 * eff + 1e-9
