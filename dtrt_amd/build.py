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


# kernels.hip is compiled once per estimator precision (DT_REAL double in namespace
# dt::p64, float in dt::p32) and, within each, once per dual width N (its AD, child,
# primary-edge, adjoint and finalize kernels; each object holds its own
# frame-rotation constant) and once without DT_N (the width-independent kernels and
# the dispatchers over nder). dispatch.cpp picks the precision, host.cpp is the C-ABI.
WIDTHS = (16, 8, 4, 3, 2, 1)   # widest first: the longest compiles start first
PRECISIONS = (("p64", "double"), ("p32", "float"))


def _jobs():
    kern = os.path.join(SRC, "kernels.hip")
    jobs = []
    for ns, real in PRECISIONS:
        prec = ["-DDT_REAL=" + real, "-DDT_PNS=" + ns]
        jobs += [(kern, "kernels_%s_n%d" % (ns, n), prec + ["-DDT_N=%d" % n, "-DDT_FRAME_ROT=c_frame_rot_n%d" % n])
                 for n in WIDTHS]
    for ns, real in PRECISIONS:
        jobs.append((kern, "kernels_" + ns, ["-DDT_REAL=" + real, "-DDT_PNS=" + ns]))
    jobs.append((os.path.join(SRC, "dispatch.cpp"), "dispatch", []))
    jobs.append((os.path.join(SRC, "scratch.hip"), "scratch", []))
    jobs.append((os.path.join(SRC, "host.cpp"), "host", []))
    return jobs


def _compile(job):
    src, name, defs = job
    obj = os.path.join(OBJ, name + ".o")
    cmd = [HIPCC] + FLAGS + defs + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):   # host code: HIP runtime API only, no device pass
        cmd[1:1] = ["-x", "hip", "--offload-host-only"]
    subprocess.run(cmd, check=True)
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = _jobs()
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, os.path.getsize(LIB), "bytes")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
