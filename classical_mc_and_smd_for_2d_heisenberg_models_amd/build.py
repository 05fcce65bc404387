"""Builds libhsmd.so (HIP kernels + C ABI) in-tree for gfx950.

Usage: python -m classical_mc_and_smd_for_2d_heisenberg_models_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhsmd.so")
BUILD = os.path.join(HERE, "_build")

SOURCES = ["lattice.cpp", "points.cpp", "api.cpp", "mc.hip", "sq.hip", "smd.hip", "smd_f32.hip", "smd_f64.hip", "dssf_mfma.hip", "dssf_fft.hip", "dssf_direct.hip", "driver.cpp", "postproc.cpp", "tune.cpp", "qc.cpp", "stats.cpp", "noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-I" + os.path.join(HERE, "..", "include"),
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    # the drop-in executables (xMC_run / xSMD_run) over the C ABI
    for exe in ("xMC_run", "xSMD_run"):
        dst = os.path.join(HERE, exe)
        cmd = ["g++", "-O2", "-std=c++17", os.path.join(CSRC, "driver_main.cpp"), "-o", dst + ".tmp",
               "-L" + HERE, "-lhsmd", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("driver link failed: %s\n%s" % (r.stdout, r.stderr))
        os.replace(dst + ".tmp", dst)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
