"""Build recipe of liboercel_mi.so (gfx950) in-tree.

    python welcome-youtube_amd/build.py      # or __graft_entry__.build()

FP contraction is off for the whole library: the FP64 parts (integration,
cell keys, chemistry distances, bonded forces) must produce the same bits as
the CPU restatement; the FP32 pair arithmetic is written in the operation
order it needs.  Packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "oc_capi.hip")
OUT = os.path.join(HERE, "liboercel_mi.so")
ARCH = os.environ.get("OC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose=False, extra=()):
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + FLAGS + list(extra) + ["-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    if verbose and (r.stdout or r.stderr):
        print(r.stdout + r.stderr, file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
