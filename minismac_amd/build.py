"""Build libsmac_amd.so (hipcc, gfx950 only) in-tree next to this file."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsmac_amd.so")
SOURCES = ["api.hip", "assemble.hip", "linerelax.hip", "lusgs.hip", "sa.hip", "forces.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: the kernels keep the reference's operation order (no FMA contraction),
# which is what makes the FP64 path bit-identical to the CPU restatement.
# -packed-fp32-ops: no v_pk_*_f32 in any kernel.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
         "-Wno-unused-function", "-Wno-unused-variable"]


def build_library(verbose: bool = False) -> str:
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops", "is not a recognized feature", "not a recognized feature")
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if not any(n in l for n in noise)]
    if r.returncode != 0:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError("hipcc failed building libsmac_amd.so")
    if verbose and msgs:
        sys.stderr.write("\n".join(msgs) + "\n")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
