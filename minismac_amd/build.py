"""Build libsmac_amd.so (hipcc, gfx950 only) in-tree next to this file.

Each source is compiled to its own object in parallel (objects under _build/, git-ignored),
then the objects are linked into the shared library."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libsmac_amd.so")
SOURCES = ["api.hip", "assemble.hip", "linerelax.hip", "linescan.hip", "lusgs.hip", "sa.hip",
           "forces.hip", "noalloc.hip", "adjoint.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: the kernels keep the reference's operation order (no FMA contraction),
# which is what makes the FP64 path bit-identical to the CPU restatement.
# -packed-fp32-ops: no v_pk_*_f32 in any kernel.
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
         "-Wno-unused-function", "-Wno-unused-variable"]
NOISE = ("packed-fp32-ops", "is not a recognized feature", "not a recognized feature")


def _filter(text: str) -> list[str]:
    return [l for l in text.splitlines() if not any(n in l for n in NOISE)]


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    for s in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", obj]
        procs.append((s, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                               stderr=subprocess.STDOUT, text=True)))
    failed, objs = [], []
    for s, obj, p in procs:
        out, _ = p.communicate()
        msgs = _filter(out)
        if p.returncode != 0:
            failed.append(s)
            sys.stderr.write(f"--- {s}\n" + "\n".join(msgs) + "\n")
        elif verbose and msgs:
            sys.stderr.write(f"--- {s}\n" + "\n".join(msgs) + "\n")
        objs.append(obj)
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write("\n".join(_filter(r.stdout + r.stderr)) + "\n")
        raise RuntimeError("linking libsmac_amd.so failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
