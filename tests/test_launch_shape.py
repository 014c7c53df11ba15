"""Launch shapes of the tiled stencils (CPU): every z-chunking a plan admits
fits the monitor partials the plan allocates.

The tiled kernels write one monitor partial per workgroup and slot at
slot * pblocks + block; a launch with more workgroups than pblocks writes
past the partials buffer.  nfgs_amd/csrc/launch_shape.h holds the shape
arithmetic that the launchers, the plan's partials sizing and
nfgs_plan_set_tuning share; tests/native/launch_shape_check.cpp checks it with
the host compiler over a sweep of grids and CU counts, including the grids of
the batched Robin-weight scan (130, 258 and 386 columns over 129 planes),
where tuned chunk counts launch up to six times the cost model's blocks.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_every_admissible_chunking_fits_the_partials(tmp_path):
    exe = tmp_path / "launch_shape_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror",
                    "-I", str(ROOT / "nfgs_amd" / "csrc"),
                    str(ROOT / "tests" / "native" / "launch_shape_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok "), out.stdout + out.stderr
    assert int(out.stdout.split()[1]) > 100000
