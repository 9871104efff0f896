"""Builds libchroma_mi355x.so (HIP for gfx950) in-tree: ``python -m chroma_amd.build [--incremental]``."""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchroma_mi355x.so")

SOURCES = ["hop_p64.hip", "hop_p32.hip", "hop_p16.hip", "hop_multi.hip", "hop_mfma.hip", "blas.hip", "setup.hip", "krylov.hip", "learned.hip",
           "op.cpp", "batch.cpp", "solvers.cpp", "multishift.cpp", "deriv.hip", "force.hip", "stout.hip", "meson.hip", "mg.hip", "mg.cpp", "fgmres.cpp", "learned.cpp", "capi.cpp", "gauge_gen.cpp", "qio.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
         # packed FP32 instructions are not emitted (see the build notes of this project)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wall", "-Wno-unused-function",
         "-Wno-unused-variable", "-Wno-unused-result"]


def _deps_newer(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    for d in [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")] + \
             [os.path.join(ROOT, "include", "chroma_mi355x.h")]:
        if os.path.getmtime(d) > t:
            return True
    return False


def _compile(src, incremental):
    s = os.path.join(CSRC, src)
    o = os.path.join(BUILD, src + ".o")
    if incremental and not _deps_newer(s, o):
        return src, 0, ""
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", s, "-o", o]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return src, p.returncode, p.stdout + p.stderr


def build(incremental=False, jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)), 16)
    errors = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, out in ex.map(lambda s: _compile(s, incremental), SOURCES):
            if rc != 0:
                errors.append((src, out))
            elif verbose and out.strip():
                print(out)
    if errors:
        msg = "\n".join(f"== {s} ==\n{o}" for s, o in errors)
        raise RuntimeError("HIP build failed:\n" + msg)
    objs = [os.path.join(BUILD, s + ".o") for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    build_host()
    return LIB


HOST = os.path.join(HERE, "host")
HOST_PROP = os.path.join(HOST, "mi355x_prop")


def build_host():
    """The C++ host layer (host/mi355x_solver.{h,cpp}) and its propagator driver, with g++ against the
    library; the driver finds the library through its run path."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", HOST,
           os.path.join(HOST, "mi355x_solver.cpp"), os.path.join(HOST, "mi355x_prop.cpp"),
           "-L", HERE, "-lchroma_mi355x", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib",
           "-o", HOST_PROP + ".tmp"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("host build failed:\n" + p.stdout + p.stderr)
    os.replace(HOST_PROP + ".tmp", HOST_PROP)
    return HOST_PROP


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(incremental=a.incremental, verbose=a.verbose))
    sys.exit(0)
