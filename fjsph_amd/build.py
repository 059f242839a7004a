"""Build the engine library fjsph_amd/libfjsph_engine.so for gfx950, and the
restart-file library fjsph_amd/libfjsph_restart.so.

hipcc compiles the kernels and the host C-ABI in one shared object.  Packed
FP32 instructions are disabled (-packed-fp32-ops) on purpose; the FP32 paths
are scalar.  The restart library is host code (g++) over the engine's C-ABI
and the HDF5 C library of the image (/opt/conda).  The CPU restatement used
as the test oracle is built separately by oracle/Makefile.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfjsph_engine.so")
RESTART_LIB = os.path.join(HERE, "libfjsph_restart.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HDF5_ROOT = os.environ.get("FJSPH_HDF5_ROOT", "/opt/conda")
SYSTEM_LIBDIR = "/usr/lib/x86_64-linux-gnu"
SOURCES = ["kernels.hip", "pairs.hip", "mesh.hip", "surrogate.hip", "surrogate_train.hip", "inlet.hip", "ipt.hip",
           "noalloc.hip", "engine.cpp", "settings.cpp", "tau.cpp"]
NO_CONTRACT = {"settings.cpp", "ipt.hip"}  # operation order of the CPU restatement, no fused multiply-adds
HOST_ONLY = {"settings.cpp", "tau.cpp"}  # plain C++, no device code
ARCH = os.environ.get("FJSPH_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, lib=LIB, build_dir=None, defines=()):
    """Compile every source (in parallel) and link `lib`.  `defines` are
    extra -D options for schedule variants."""
    from concurrent.futures import ThreadPoolExecutor

    build_dir = build_dir or os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    common = [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
    ] + [f"-D{d}" for d in defines]
    cmds, objs = [], []
    for src in SOURCES:
        obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith((".hip", ".cpp")) and src not in HOST_ONLY else []
        extra = ["-ffp-contract=off"] if src in NO_CONTRACT else []
        inc = [f"-I{INCLUDE}"]
        cmds.append([hipcc()] + common + extra + lang + inc + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for f in [ex.submit(run, c) for c in cmds]:
            f.result()
    tmp = lib + ".tmp"
    run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


def build_restart(verbose=False):
    """Link libfjsph_restart.so against libfjsph_engine.so and libhdf5."""
    inc = os.path.join(HDF5_ROOT, "include")
    libdir = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        raise RuntimeError(f"hdf5.h not found under {inc}")
    tmp = RESTART_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}", f"-I{inc}",
           os.path.join(CSRC, "restart.cpp"), "-o", tmp, f"-L{HERE}", "-lfjsph_engine", f"-L{libdir}", "-lhdf5",
           # the system library directory comes before the HDF5 one, so that
           # only libhdf5 (and not an older libstdc++) is taken from there
           f"-Wl,-rpath,$ORIGIN:{SYSTEM_LIBDIR}:{libdir}", "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, RESTART_LIB)
    return RESTART_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_restart(verbose="-v" in sys.argv))
