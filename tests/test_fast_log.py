"""The FP64 logarithm of the node records (csrc/fastlog.hpp: argument
reduction to [sqrt(1/2), sqrt(2)) and the atanh series), compiled for the host
from the same source the kernels include, against the x87 extended-precision
logl: at most 2 ulp over random positive doubles (uniform bit patterns,
subnormals, values near 1 and near the reduction boundary, the densities and
pressures of order one the rhs! sees, all powers of two), and IEEE special
values."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "trixicuda.jl_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("fastlog") / "fast_log_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "native", "fast_log_check.cpp"),
                    "-o", exe], check=True)
    return exe


def test_fast_log_within_two_ulp(checker):
    out = subprocess.run([checker, "1000000"], capture_output=True, text=True, check=True).stdout
    lines = dict(l.split(" ", 1) for l in out.strip().splitlines())
    worst = float(lines["max_ulp"].split()[0])
    print(out)
    assert worst <= 2.0, out
    assert lines["special"] == "ok", out
