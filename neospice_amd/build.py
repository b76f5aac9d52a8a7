"""Builds neospice_amd/libneospice_amd.so for gfx950 with hipcc.

All HIP code is compiled without FMA contraction (-ffp-contract=off) so the
FP64 arithmetic matches the reference CPU path operation for operation, and
with packed FP32 ops disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libneospice_amd.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = [
    "circuit.cpp",
    "devices.cpp",
    "bsim4.cpp",
    "lu_structure.cpp",
    "lu_cpu.cpp",
    "batch_lu.cpp",
    "engine.cpp",
    "adjoint.cpp",
    "small_lu.cpp",
    "bundle.cpp",
    "ac_lu.cpp",
    "capi.cpp",
    "gpu_lu.hip",
    "gpu_dev.hip",
    "gpu_nonlinear.hip",
    "gpu_bsim4.hip",
    "gpu_adjoint.hip",
    "gpu_small_lu.hip",
    "gpu_bundle.hip",
    "noalloc.hip",
    "gpu_equil.hip",
    "gpu_small_lu_c.hip",
]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    cmd = [hipcc()] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + HIP_FLAGS
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    workers = min(8, len(SOURCES))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
