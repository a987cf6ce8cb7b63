"""Builds the native libraries of classicaldft_amd in-tree.

* ``libcdft.so``   - HIP kernels + the C ABI of ``include/cdft.h`` (hipcc, gfx950)
* ``libclassicaldft.so`` - the C++ drop-in API of ``include/classicaldft/`` on top of it

Run ``python -m classicaldft_amd.build`` (or ``__graft_entry__.build()``).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
ARCH = os.environ.get("CDFT_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep the compiler from emitting packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HIP_SOURCES = ("cdft_api.hip", "pointwise.hip", "solvers.hip", "batch.hip", "unet.hip")
# host-only C++ of libcdft.so, compiled by g++ without floating-point contraction (weights_ref.cpp reproduces the
# reference's double expressions operation for operation)
HOST_SOURCES = ("weights_ref.cpp",)
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_cdft():
    """Each translation unit compiled in its own hipcc process (in parallel), then one link."""
    out = os.path.join(HERE, "libcdft.so")
    tmp = out + ".tmp"
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, f[:-4] + ".o") for f in HIP_SOURCES]
    cmds = [[hipcc(), *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", "-o", o, os.path.join(CSRC, f)]
            for f, o in zip(HIP_SOURCES, objs)]
    hobjs = [os.path.join(objdir, f[:-4] + ".o") for f in HOST_SOURCES]
    cmds += [["g++", *HOST_FLAGS, "-I", os.path.join(ROOT, "include"), "-c", "-o", o, os.path.join(CSRC, f)]
             for f, o in zip(HOST_SOURCES, hobjs)]
    objs = objs + hobjs
    for c in cmds:
        print(" ".join(c), flush=True)
    procs = [subprocess.Popen(c) for c in cmds]
    codes = [p.wait() for p in procs]
    if any(codes):
        raise subprocess.CalledProcessError(max(codes), cmds[[i for i, c in enumerate(codes) if c][0]])
    run([hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_host():
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp")) if os.path.isdir(HOST) else []
    if not srcs:
        return None
    out = os.path.join(HERE, "libclassicaldft.so")
    tmp = out + ".tmp"
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs,
         "-L", HERE, "-lcdft", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    ex = os.path.join(HOST, "examples")
    for f in sorted(os.listdir(ex)) if os.path.isdir(ex) else []:
        if f.endswith(".cpp"):
            exe = os.path.join(HERE, f[:-4])
            run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", exe + ".tmp",
                 os.path.join(ex, f), "-L", HERE, "-lclassicaldft", "-lcdft", "-Wl,-rpath,$ORIGIN"])
            os.replace(exe + ".tmp", exe)
    return out


def build_all():
    build_cdft()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
