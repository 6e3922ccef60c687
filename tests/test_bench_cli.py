"""bench.py's command line on the CPU: more than one GPU is refused with a
message and a non-zero status before anything touches a device, and the
flop and byte model it reports is the one frozen in flop_model.py."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_headline_model_figures():
    spec = importlib.util.spec_from_file_location("fm", os.path.join(ROOT, "trixicuda.jl_amd", "flop_model.py"))
    fm = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fm)
    key = "euler3d_fd_ranocha_llf"
    # read u and write du, FP64, five variables
    assert fm.BYTES_PER_DOF[key] == 80
    assert abs(fm.FLOPS_PER_DOF[key] - 405) < 1e-9
    assert fm.PEAK["hbm_GBps"] == 8000.0
