"""Build the gfx950 HIP library ``adrt_amd/libadrt_hip.so`` in-tree.

Usage: ``python -m adrt_amd.build`` (or ``__graft_entry__.build()``).

Flags:
  --offload-arch=gfx950        CDNA4 only, no other targets
  -ffp-contract=off            products are rounded before sums (high-pass)
  -packed-fp32-ops disabled    no v_pk_*_f32 generated from float code
Denormals are kept (no flush) and no fast-math is used, so each IEEE
operation matches the CPU reference's.
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libadrt_hip.so")
SOURCES = ["capi.hip", "fwd.hip", "fused.hip", "small.hip", "bdrt.hip", "iadrt.hip", "fmg.hip",
           "layout.hip", "nn.hip", "cg.hip", "interp.hip", "noalloc.hip", "dctp.hip",
           "queue.hip"]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the ROCm toolchain is required to build adrt_amd")
    return path


def flags():
    return [
        "--offload-arch=gfx950",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I" + os.path.join(ROOT, "include"),
    ]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *flags(), *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libadrt_hip.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
