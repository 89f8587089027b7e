"""Build libdrh.so (the HIP kernels + the C ABI) in-tree for gfx950 with hipcc.

    python -m diffrender_tutorials_amd.build

Flags: -ffp-contract=off keeps every float expression an unfused IEEE operation (the
reference's arithmetic; hipcc defaults to fp-contract=fast); HIP's default correctly
rounded f32 divide / sqrt stay on; -munsafe-fp-atomics lets atomicAdd(float) use the
hardware global_atomic_add_f32; packed-fp32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdrh.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DRH_OFFLOAD_ARCH", "gfx950")

COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function"]
DEVICE = [f"--offload-arch={ARCH}", "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build(verbose: bool = True) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    kobj = os.path.join(objdir, "drh_kernels.o")
    aobj = os.path.join(objdir, "drh_api.o")
    _run([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "drh_kernels.hip"), "-o", kobj])
    _run([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "drh_api.cpp"), "-o", aobj])
    sobj = os.path.join(objdir, "drh_server.o")
    _run([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, "drh_server.cpp"), "-o", sobj])
    tmp = LIB + ".tmp"
    _run([HIPCC, *DEVICE, "-shared", "-fPIC", kobj, aobj, sobj, "-o", tmp, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
