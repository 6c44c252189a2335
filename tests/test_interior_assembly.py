"""The Vanka kernels resolve the boundary tests of the local-block assembly
at compile time for tiles of interior cells (tiles_device.h vanka_cell,
INTERIOR = true).  That is exact when every interior cell (i, j) of any grid
assembles the same bits as the cell (1, 1) of a 3 x 3 grid; this runs the
device header's own assembly on the host to check it over every interior
cell of several grids, with random neighbourhoods, and checks that a boundary
cell does differ."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_interior_constants_reproduce_every_interior_cell(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "check"
    (tmp_path / "hip").mkdir()
    (tmp_path / "hip" / "hip_runtime.h").write_text("// host build: no HIP runtime needed\n")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", str(tmp_path),
                    "-I", os.path.join(ROOT, "glide_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    os.path.join(HERE, "interior", "check.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")
    assert int(out.stdout.split()[1]) > 1000
