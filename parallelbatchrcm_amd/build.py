"""Build the gfx950 engine library in-tree: parallelbatchrcm_amd/libpbrcm.so.

One hipcc invocation compiles the HIP engine (csrc/*.hip) and the host C++ workload generators
(csrc/*.cpp) into a single shared library exporting the C ABI of include/rcm_gpu_batch_abi.h and
include/rcm_workloads.h.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpbrcm.so")
ARCH = os.environ.get("PBRCM_ARCH", "gfx950")


def sources() -> list[str]:
    csrc = os.path.join(HERE, "csrc")
    return sorted(
        os.path.join(csrc, f) for f in os.listdir(csrc) if f.endswith((".hip", ".cpp"))
    )


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [
        hipcc,
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        # no packed-fp32 instructions anywhere in device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        "-I", os.path.join(HERE, "csrc"),
        "-pthread",
        "-o", LIB + ".tmp",
        *sources(),
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(
        line for line in (res.stderr or "").splitlines() if "packed-fp32-ops" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
