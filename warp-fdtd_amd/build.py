"""Build libwfdtd.so for gfx950 with hipcc (run as a script or via build())."""

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = ["yee.hip", "detect.hip", "adjoint.hip", "driven.hip", "material.hip",
           "stage.hip", "resident.hip", "design.hip", "fno.hip", "runtime.hip"]
# Every float operation of the kernels rounds on its own (no FMA contraction),
# so they reproduce the reference's operation order bit for bit; packed FP32
# instructions are disabled.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall"]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    out = HERE / "libwfdtd.so"
    tmp = HERE / "libwfdtd.so.tmp"
    cmd = [hipcc, *FLAGS, *[str(HERE / "csrc" / s) for s in SOURCES], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("hipcc failed building libwfdtd.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
