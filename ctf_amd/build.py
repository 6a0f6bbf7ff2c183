"""Builds the native library ``ctf_amd/libmaragpu.so`` for gfx950 with hipcc, and the
checkpoint library ``ctf_amd/libmara_h5.so`` with g++.

libmaragpu.so holds the HIP kernels (``csrc/*.hip``), the C ABI of ``include/maragpu.h``
and the Lua-verb host layer of ``include/mara_host.h``. libmara_h5.so holds the HDF5
checkpoint files of ``include/mara_checkpoint.h``; it links the system's HDF5 (found
through ``HDF5_DIR``, else /opt/conda, else /usr) and is skipped when there is none. Both
are built in-tree so the shared objects travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmaragpu.so")
H5LIB = os.path.join(HERE, "libmara_h5.so")

SOURCES = ["maragpu.hip", "mlp.hip", "mara_app.cpp", "initial_data.cpp", "ou_field.cpp"]
HEADERS = ["rmhd_device.hpp", "mlp_device.hpp", "eigensystem_device.hpp", "ou_field.hpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
    "-Wno-unused-value", "-Wno-unused-result",
    # packed FP32 instructions off: some of their forms have given wrong results, without an
    # error, on these GPUs
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core accumulators in ordinary VGPRs: the network's activations are read by
    # VALU right after each product, which the accumulator-register form makes a copy
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wl,-soname,libmaragpu.so",
]


def hdf5_dir() -> str | None:
    """The prefix holding include/hdf5.h and lib/libhdf5.so, or None."""
    for d in (os.environ.get("HDF5_DIR"), "/opt/conda", "/usr"):
        if d and os.path.exists(os.path.join(d, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(d, "lib", "libhdf5.so")):
            return d
    return None


def build_h5(verbose: bool = False) -> str | None:
    """Compiles csrc/checkpoint.cpp into libmara_h5.so against HDF5 and libmaragpu.so.
    Returns its path, or None when no HDF5 installation is found."""
    h5 = hdf5_dir()
    if h5 is None:
        sys.stderr.write("[build] no HDF5 installation found: checkpoint library not built\n")
        return None
    tmp = H5LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(h5, "include"),
           "-o", tmp, os.path.join(CSRC, "checkpoint.cpp"),
           "-L", os.path.join(h5, "lib"), "-lhdf5", "-Wl,-rpath," + os.path.join(h5, "lib"),
           "-L", HERE, "-lmaragpu", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building libmara_h5.so")
    os.replace(tmp, H5LIB)
    return H5LIB


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def build(verbose: bool = False) -> str:
    """Compiles every source into libmaragpu.so (always recompiles). Returns its path."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libmaragpu.so")
    os.replace(tmp, LIB)
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
