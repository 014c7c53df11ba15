"""Bottom-plane re-separation error E^rel_{B^>} (nfgs_amd.cci) on the GPU.

The reference helpers (nfgs/_backend/scripts/plot_nfgs_td_robin_metrics_vs_er.py:240-326)
split each field's z = 0 plane with an unpadded NFGS instance and average
|b^> - B^>| / |b^>|.  Pinned here by closed forms on boundaries built from
potential fields of sources above and below the plane (the outward part of a
sum is the part of the sources above; tests/test_oracle_above_field.py), by
linearity (a solution twice the reference scores exactly 1), and against the
same mean over the oracle's outward fields.
"""
import sys

import numpy as np
import pytest
import torch

from oracle.nfgs_oracle import OracleNFGS
from tests.golden_cases import rel_l2
from tests.test_oracle_above_field import DX, DY, NX, NY, _modes, _potential_boundary

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


def _fields(bx, by, bz, nz=3):
    # a (nz, ny, nx) volume whose bottom plane is the boundary; the planes
    # above are arbitrary (the split reads z = 0 only)
    rng = np.random.default_rng(11)
    out = {}
    for k, v in (("Bx", bx), ("By", by), ("Bz", bz)):
        vol = rng.standard_normal((nz,) + v.shape)
        vol[0] = v
        out[k] = vol
    return out


def _helper():
    from nfgs_amd.cci import make_bgt_helper_solver
    return make_bgt_helper_solver(nx=NX, ny=NY, dx=DX, dy=DY, dz=1.0)


def test_outward_part_ignores_the_sources_below():
    _announce("cci closed forms")
    from nfgs_amd.cci import bottom_bgt_relative_error, bottom_plane_bgt_nfgs
    up = _potential_boundary(_modes(5), -1.0)
    down = _potential_boundary(_modes(6), 1.0)
    s = _helper()
    ref = _fields(*(u + d for u, d in zip(up, down)))
    # the split of the sum is the part of the sources above
    bgt = bottom_plane_bgt_nfgs(s, ref)
    assert bgt.shape == (3, NY, NX)
    assert rel_l2(bgt, np.stack(up)) < 1e-12
    # changing only the sources below leaves B^> as it was
    sol = _fields(*(u + 2.0 * d for u, d in zip(up, down)))
    assert bottom_bgt_relative_error(s, ref, sol) < 1e-11
    # a solution without the sources above has no outward part: every pixel scores 1
    assert bottom_bgt_relative_error(s, ref, _fields(*down)) == pytest.approx(1.0, abs=1e-11)
    # the split is linear: twice the reference scores exactly 1
    twice = {k: 2.0 * v for k, v in ref.items()}
    assert bottom_bgt_relative_error(s, ref, twice) == pytest.approx(1.0, abs=1e-11)
    assert bottom_bgt_relative_error(s, ref, ref) < 1e-14


def test_matches_the_mean_over_oracle_splits():
    _announce("cci vs oracle")
    from nfgs_amd.cci import bottom_bgt_relative_error
    rng = np.random.default_rng(8)
    ref = _fields(*(rng.standard_normal((NY, NX)) for _ in range(3)))
    sol = {k: v + 0.2 * rng.standard_normal(v.shape) for k, v in ref.items()}

    def oracle_bgt(f):
        o = OracleNFGS(NX, NY, 8, DX, DY, 1.0, dtype=np.float64, pad_calculation_domain=False,
                       source_zero_mode_policy="project", d0_zero_mode_policy="project")
        o.set_bottom_vector_boundary(f["Bx"][0], f["By"][0], f["Bz"][0])
        return o.gaussian_above_bottom_field()

    b, B = oracle_bgt(ref), oracle_bgt(sol)
    ys, xs = slice(2, NY - 3), slice(1, NX - 2)
    for sy, sx in ((slice(None), slice(None)), (ys, xs)):
        bm = np.linalg.norm(b[:, sy, sx], axis=0)
        dm = np.linalg.norm((b - B)[:, sy, sx], axis=0)
        mask = bm > 1e-8 * bm.max()
        want = float(np.mean(dm[mask] / bm[mask]))
        got = bottom_bgt_relative_error(_helper(), ref, sol, y_slice=sy, x_slice=sx)
        assert got == pytest.approx(want, rel=1e-10)


def test_no_valid_pixels_raises():
    _announce("cci empty mask")
    from nfgs_amd.cci import bottom_bgt_relative_error
    zero = _fields(*(np.zeros((NY, NX)) for _ in range(3)))
    with pytest.raises(RuntimeError, match="No valid bottom pixels"):
        bottom_bgt_relative_error(_helper(), zero, zero)
