"""Pins the algorithmic flop counts used for the roofline (glide_amd/roofline.py)
by running the solver's own per-cell code on the host with a counting
scalar type.  The device headers are copied with `double` replaced by the
counting type R; nothing else in them changes."""
import os
import re
import shutil
import subprocess

import pytest

from glide_amd import roofline

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "glide_amd", "csrc")


def _transform(src: str) -> str:
    s = src.replace("#include <hip/hip_runtime.h>", "")
    s = re.sub(r"\bdouble\b", "R", s)
    # bit-level helpers of detmath keep real doubles
    s = s.replace("typedef union { R d; uint64_t u; } dm_bits_t;",
                  "typedef union { double d; uint64_t u; } dm_bits_t;")
    s = s.replace("DM_FN R dm_from_bits(uint64_t u) { dm_bits_t b; b.u = u; return b.d; }",
                  "DM_FN R dm_from_bits(uint64_t u) { dm_bits_t b; b.u = u; return R(b.d); }")
    s = s.replace("DM_FN uint64_t dm_to_bits(R d) { dm_bits_t b; b.d = d; return b.u; }",
                  "DM_FN uint64_t dm_to_bits(R d) { dm_bits_t b; b.d = (double)d; return b.u; }")
    s = s.replace("__builtin_floor(x)", "count_floor(x)")
    s = s.replace("(R)(nyc", "R((double)(nyc").replace("(R)e;", "R((double)e);")
    return s


@pytest.fixture(scope="module")
def counts(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("flopcount")
    for h in ("detmath.h", "ssa_device.h", "vanka_device.h"):
        with open(os.path.join(CSRC, h)) as f:
            (d / h).write_text(_transform(f.read()))
    exe = d / "count"
    subprocess.run(["g++", "-std=c++17", "-O0", "-I", str(d), "-I", os.path.join(ROOT, "include"),
                    "-o", str(exe), os.path.join(HERE, "flopcount", "count.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    return int(out[0]), int(out[1]), int(out[2]), int(out[3])


def test_newton_iteration_flops(counts):
    newton, _, constrained, _ = counts
    assert constrained == 0
    assert newton == roofline.FLOPS_NEWTON_ITER, newton


def test_eta_flops(counts):
    _, eta, _, _ = counts
    assert eta == roofline.FLOPS_ETA, eta


def test_cell_setup_flops(counts):
    _, _, _, setup = counts
    assert setup == roofline.FLOPS_CELL_SETUP, setup
