"""Build step of the package: compiles the HIP library for gfx950 in-tree.

    python -m 2ddrugsimulation_amd.build      (or __graft_entry__.build())
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed-FP32 instructions are disabled for all device code
FLAGS = ["-O3", "-fPIC", "-shared", "-std=c++17", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def build_library(verbose: bool = False) -> str:
    src = os.path.join(PKG_DIR, "csrc", "cep_gpu.hip")
    out = os.path.join(PKG_DIR, "libcepgpu.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libcepgpu.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
