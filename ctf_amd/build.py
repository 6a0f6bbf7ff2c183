"""Builds the native library ``ctf_amd/libmaragpu.so`` for gfx950 with hipcc.

The library holds the HIP kernels (``csrc/*.hip``), the C ABI of ``include/maragpu.h``
and the Lua-verb host layer of ``include/mara_host.h``. It is built in-tree so the
shared object travels with the source tree.
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

SOURCES = ["maragpu.hip", "mlp.hip", "mara_app.cpp", "initial_data.cpp", "ou_field.cpp"]
HEADERS = ["rmhd_device.hpp", "mlp_device.hpp", "eigensystem_device.hpp", "ou_field.hpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
    "-Wno-unused-value", "-Wno-unused-result",
    # packed FP32 instructions off (see DESIGN: hardware trap on this pool)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core accumulators in ordinary VGPRs: the network's activations are read by
    # VALU right after each product, which the accumulator-register form makes a copy
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]


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
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
