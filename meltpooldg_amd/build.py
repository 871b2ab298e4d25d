"""Builds the HIP library libmpdg_heat.so in-tree for gfx950 with hipcc.

Every HIP source is compiled with packed FP32 instructions disabled and with
hardware floating-point atomics; the objects are linked into one shared
library next to this file, which the Python mirror loads with ctypes.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpdg_heat.so")
SOURCES = ["heat_apply.hip", "heat_aux.hip", "blas.hip", "gmg_kernels.hip", "tangent.hip", "cnn.hip", "heat_lib.cpp",
           "learned.cpp", "io.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         "-Wno-unused-variable", "-Wno-unused-but-set-variable"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return obj


def build(jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-6000:])
    os.replace(tmp, LIB)
    return LIB


def build_abi_smoke():
    """Compiles tests/cxx/abi_smoke.cpp, a C++ caller of include/mpdg_heat.h, against the built
    library (the header compiled as C++ by a program that is not this package)."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cxx", "abi_smoke.cpp")
    out = os.path.join(root, "tests", "cxx", "abi_smoke")
    if not os.path.exists(src):
        return None
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-x", "hip", src, "-o", out + ".tmp",
           "-L" + HERE, "-lmpdg_heat", "-Wl,-rpath,$ORIGIN/../../meltpooldg_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("abi_smoke build failed:\n" + r.stderr[-4000:])
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build())
    sys.exit(0)
