"""Host-side helpers that need no GPU: argument and box checks of the scoring,
evaluation and NPZ modules (the arithmetic itself is tested on the GPU)."""
import numpy as np
import pytest

from nfgs_amd import controlled, reconstruct_npz, scores


def test_scoring_box_and_spacing_checks():
    assert scores._box((slice(1, 4), slice(None), slice(2, -1)), (5, 6, 7)) == [1, 4, 0, 6, 2, 6]
    with pytest.raises(ValueError):
        scores._box((slice(0, 5, 2), slice(None), slice(None)), (5, 6, 7))
    with pytest.raises(ValueError):
        scores._box((slice(3, 3), slice(None), slice(None)), (5, 6, 7))
    assert scores._spacing_xyz(0.5) == (0.5, 0.5, 0.5)
    assert scores._spacing_xyz((1.0, 2.0, 3.0)) == (1.0, 2.0, 3.0)
    for bad in (0.0, (1.0, -1.0, 1.0), (1.0, 2.0)):
        with pytest.raises(ValueError):
            scores._spacing_xyz(bad)


def test_evaluation_scoring_box():
    cfg = {"scoring_slices_zyx": [[0, 5, 1], [1, 6], [2, 8, None]]}
    box = controlled.ScoringBox.parse(cfg["scoring_slices_zyx"], (5, 7, 9))
    assert box.slices == (slice(0, 5, 1), slice(1, 6), slice(2, 8, None))
    assert box.spec == cfg["scoring_slices_zyx"] and box.window == (slice(1, 6), slice(2, 8, None))
    for bad in ([[0, 6, 1], [0, 7, 1], [0, 9, 1]], [[0, 5, 2], [0, 7, 1], [0, 9, 1]],
                [[0, 5, 1], [0, 7, 1]]):
        with pytest.raises(ValueError):
            controlled.ScoringBox.parse(bad, (5, 7, 9))


def test_npz_boundary_reader_and_arguments(tmp_path):
    np.savez(tmp_path / "b.npz", Bx=np.zeros((4, 5)), By=np.ones((4, 5)), Bz=np.ones((4, 5)),
             dx=0.5, dy=0.25, dz=1.0, Nz=6)
    bx, by, bz, nz, h = reconstruct_npz.read_boundary(tmp_path / "b.npz")
    assert bz.shape == (4, 5) and nz == 6 and h == (0.5, 0.25, 1.0)
    np.savez(tmp_path / "bad.npz", Bx=np.zeros((4, 5)), By=np.zeros((4, 5)), Bz=np.zeros((5, 4)),
             dx=1.0, dy=1.0, dz=1.0, Nz=4)
    with pytest.raises(ValueError, match="one shape"):
        reconstruct_npz.read_boundary(tmp_path / "bad.npz")
    # argument errors end before any device work
    for argv in ([str(tmp_path / "b.npz"), "--rho", "1.5"],
                 [str(tmp_path / "b.npz"), "--device", "cpu"]):
        with pytest.raises(SystemExit):
            reconstruct_npz.main(argv)
