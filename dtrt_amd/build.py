"""Build recipe of libdtrt_mi355x.so (gfx950) — in-tree, with hipcc.

Usage: ``python -m dtrt_amd.build`` or ``dtrt_amd.build.build()``. Objects are
compiled in parallel into dtrt_amd/_build/ and linked into dtrt_amd/lib/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libdtrt_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DTRT_ARCH", "gfx950")

# -packed-fp32-ops: keep hipcc from forming packed fp32 (v_pk_*_f32) instructions,
# some forms of which have returned wrong results on gfx950 without any error.
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include")]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):   # host code: HIP runtime API only, no device pass
        cmd[1:1] = ["-x", "hip", "--offload-host-only"]
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(SRC, f) for f in ("kernels.hip", "host.cpp")]
    with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, os.path.getsize(LIB), "bytes")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
