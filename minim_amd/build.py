"""Build libminim_amd.so (HIP kernels + C++ host library + C ABI) for gfx950.

Usage: python -m minim_amd.build   (or __graft_entry__.build())

Every source under minim_amd/csrc is compiled with hipcc into minim_amd/_build/
and linked into minim_amd/libminim_amd.so, next to this file, so the built
library travels with the source tree.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libminim_amd.so")
ARCH = os.environ.get("MINIM_OFFLOAD_ARCH", "gfx950")

CFLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # every product rounded before it is added, as in the reference's vector operators
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-pass-failed", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(cc, src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [cc] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(jobs=None, verbose=True):
    cc = hipcc()
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 16)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(cc, s), srcs))
    tmp = LIB + ".tmp"
    # -Bsymbolic: the library's own references to its C++ symbols bind to its
    # own definitions, so a program that also loads another library defining
    # namespace minim (the reference, as the reference-side binding does) cannot
    # interpose that library's State, Potential, ... into this one
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-Wl,-Bsymbolic", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
