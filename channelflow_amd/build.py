"""Build recipe of libcfgpu.so (the HIP kernels + C ABI), in-tree, for gfx950.

Usage: python -m channelflow_amd.build   (or via __graft_entry__.build()).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcfgpu.so")

SOURCES = ["cfg_kernels.hip", "cfg_step.cpp", "cfg_runtime.cpp", "cfg_batch.cpp", "cfg_noalloc.cpp",
           "cfg_tausolver.cpp"]
HEADERS = ["cfg_kernels.hpp", "cfg_math.hpp", "cfg_wave.hpp", "cfg_host.hpp", "cfg_step.hpp"]

# Packed FP32 instructions are disabled in every HIP compile of this project.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    """Compile the HIP sources (object per source, then link) into LIB."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    procs = []
    for src in SOURCES:  # one compiler process per source, in parallel
        obj = os.path.join(build_dir, src + ".o")
        cmd = [hipcc()] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError(f"hipcc failed on {failed}")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


H5_SOURCE = os.path.join(CSRC, "cfg_h5io.cpp")
H5_LIB = os.path.join(HERE, "libcfgh5.so")
HDF5_PREFIXES = ("/opt/conda", "/usr", "/usr/local")


def build_h5io(verbose=False):
    """libcfgh5.so: channelflow's .h5/.nc file layouts over the HDF5 C library
    (host code, g++). Returns the library path, or None when no HDF5
    installation with headers is found."""
    for pre in HDF5_PREFIXES:
        inc, lib = os.path.join(pre, "include"), os.path.join(pre, "lib")
        if os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(lib, "libhdf5.so")):
            tmp = H5_LIB + ".tmp"
            cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + inc, H5_SOURCE, "-o", tmp,
                   "-L" + lib, "-lhdf5", "-Wl,-rpath," + lib]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
            os.replace(tmp, H5_LIB)
            return H5_LIB
    return None


DRIVER_SOURCE = os.path.join(ROOT, "tests", "cxx", "dns_driver.cpp")
DRIVER = os.path.join(ROOT, "tests", "cxx", "dns_driver.bin")


def build_driver(verbose=False):
    """tests/cxx/dns_driver.bin: a C++ caller compiled with g++ against
    include/cfgpu.hpp and libcfgpu.so (the compiled side of the boundary)."""
    tmp = DRIVER + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
           DRIVER_SOURCE, "-L" + HERE, "-lcfgpu", "-Wl,-rpath,$ORIGIN/../../channelflow_amd", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_h5io(verbose="-v" in sys.argv))
    print(build_driver(verbose="-v" in sys.argv))
