"""Build libdelta_ray_amd.so (the HIP kernels + C-ABI) in-tree for gfx950.

    python -m delta_ray_amd.build_lib        # or delta_ray_amd.build_lib.build()

Compiles csrc/render.hip, csrc/mlp.hip, csrc/refit.hip and csrc/lbvh.hip with hipcc,
csrc/scene_build.cpp, csrc/load_serialized.cpp and csrc/ltc_fit.cpp with the
host C++ compiler
and csrc/ltc_blob.S (the embedded LTC table), then links the shared library
next to this file (with zlib, for the .serialized reader).  Everything is compiled with -ffp-contract=off so
that ray/triangle arithmetic rounds exactly like the CPU oracle's.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdelta_ray_amd.so")
LTC = os.path.join(HERE, "data", "ltc_tabM.f32")

ARCH = os.environ.get("DR_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-variable"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             "-Wno-unused-function"]


def _hipcc():
    for c in (os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc"), shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: %s" % " ".join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    cxx = os.environ.get("CXX", "g++")
    objs = {
        "render": os.path.join(BUILD, "render.o"),
        "mlp": os.path.join(BUILD, "mlp.o"),
        "refit": os.path.join(BUILD, "refit.o"),
        "lbvh": os.path.join(BUILD, "lbvh.o"),
        "scene": os.path.join(BUILD, "scene_build.o"),
        "serialized": os.path.join(BUILD, "load_serialized.o"),
        "ltc_fit": os.path.join(BUILD, "ltc_fit.o"),
        "ltc": os.path.join(BUILD, "ltc_blob.o"),
    }
    steps = [
        [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, "render.hip"), "-o", objs["render"]],
        [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, "mlp.hip"), "-o", objs["mlp"]],
        [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, "refit.hip"), "-o", objs["refit"]],
        [hipcc] + HIP_FLAGS + ["-c", os.path.join(CSRC, "lbvh.hip"), "-o", objs["lbvh"]],
        [cxx] + CXX_FLAGS + ["-c", os.path.join(CSRC, "scene_build.cpp"), "-o", objs["scene"]],
        [cxx] + CXX_FLAGS + ["-c", os.path.join(CSRC, "load_serialized.cpp"), "-o", objs["serialized"]],
        [cxx] + CXX_FLAGS + ["-pthread", "-c", os.path.join(CSRC, "ltc_fit.cpp"), "-o", objs["ltc_fit"]],
        ["gcc", "-c", "-DLTC_FILE=\"%s\"" % LTC, os.path.join(CSRC, "ltc_blob.S"), "-o", objs["ltc"]],
    ]
    with ThreadPoolExecutor(max_workers=7) as ex:
        outs = list(ex.map(_run, steps))
    if verbose:
        for o in outs:
            sys.stdout.write(o)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp,
          objs["render"], objs["mlp"], objs["refit"], objs["lbvh"], objs["scene"], objs["serialized"], objs["ltc_fit"], objs["ltc"], "-lz", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
