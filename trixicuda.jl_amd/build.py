"""Build libtrixihip.so (gfx950) in-tree with hipcc.

    python trixicuda.jl_amd/build.py [--jobs N]

Every *.hip under csrc/ is compiled to csrc/../_build/*.o in parallel and
linked into trixicuda.jl_amd/libtrixihip.so.  Packed-FP32 instructions are
disabled for all device code.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtrixihip.so")
ARCH = os.environ.get("TRIXI_HIP_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-pass-failed",
]


def _compile(src, build_dir=BUILD, extra=()):
    obj = os.path.join(build_dir, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC] + CFLAGS + list(extra) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "not a recognized feature" not in l
                     and "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msgs}")
    return obj, msgs


def build(jobs=None, verbose=False, out=LIB, build_dir=BUILD, defines=()):
    """Compile and link; `defines` are extra -D options (tuning variants)."""
    os.makedirs(build_dir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(16, os.cpu_count() or 4)
    extra = [f"-D{d}" for d in defines]
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, msgs in ex.map(lambda s: _compile(s, build_dir, extra), srcs):
            objs.append(obj)
            if verbose and msgs.strip():
                print(msgs, file=sys.stderr)
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    return out


H5_INCLUDE = os.environ.get("TRIXI_H5_INCLUDE", "/opt/conda/include")
H5_LIBDIR = os.environ.get("TRIXI_H5_LIBDIR", "/opt/conda/lib")
H5_LIB = os.path.join(HERE, "libtrixih5.so")


def build_h5(out=H5_LIB):
    """Host-only HDF5 solution writer (csrc/trixi_h5.cpp) against the HDF5 C
    library; returns None, building nothing, when its header is absent."""
    if not os.path.exists(os.path.join(H5_INCLUDE, "hdf5.h")):
        return None
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{H5_INCLUDE}",
           os.path.join(CSRC, "trixi_h5.cpp"), "-o", tmp, f"-L{H5_LIBDIR}", "-lhdf5",
           f"-Wl,-rpath,{H5_LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for trixi_h5.cpp:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
    print(build_h5())
