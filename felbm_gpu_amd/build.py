"""Build the engine library in-tree for gfx950 (hipcc) and the host pieces.

Run as ``python -m felbm_gpu_amd.build`` or through ``__graft_entry__.build()``.
The product library is ``felbm_gpu_amd/libfelbm_mi355x.so``; its C ABI is
declared in ``include/felbm_mi355x.h``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfelbm_mi355x.so")
H5LIB = os.path.join(HERE, "libfelbm_h5.so")
# HDF5 1.10 C library for the output/restart files (libfelbm_h5.so only)
HDF5_DIR = os.environ.get("FELBM_HDF5_DIR", "/opt/conda")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "--offload-arch=gfx950"
# packed FP32 instructions are disabled for all device code
NOPK = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", ARCH, "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result"] + NOPK

UNITS = [
    # (source, extra flags)
    ("engine_f64.hip", ["-ffp-contract=off"]),
    ("engine_f32.hip", ["-ffp-contract=fast"]),
    ("capi.cpp", ["-x", "hip", "-ffp-contract=off"]),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(BUILD, src + ".o")
        _run([HIPCC] + COMMON + extra + ["-c", "-o", obj, os.path.join(CSRC, src)])
        objs.append(obj)
    gobj = os.path.join(BUILD, "geometry.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-c", "-o", gobj,
          os.path.join(CSRC, "geometry.cpp")])
    objs.append(gobj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_h5(verbose)
    return LIB


def build_h5(verbose=False):
    """libfelbm_h5.so: the HDF5 writers, linked against HDF5_DIR's libhdf5."""
    if not os.path.exists(os.path.join(HDF5_DIR, "include", "hdf5.h")):
        sys.stderr.write("HDF5 headers not found under %s: libfelbm_h5.so not built\n" % HDF5_DIR)
        return None
    tmp = H5LIB + ".tmp"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(HDF5_DIR, "include"), "-o", tmp, os.path.join(CSRC, "h5io.c"),
          "-L" + os.path.join(HDF5_DIR, "lib"), "-lhdf5", "-Wl,-rpath," + os.path.join(HDF5_DIR, "lib")])
    os.replace(tmp, H5LIB)
    if verbose:
        print("built", H5LIB)
    return H5LIB


if __name__ == "__main__":
    build(verbose=True)
