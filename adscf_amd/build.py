"""Builds libadscf_hip.so for gfx950 with hipcc (in-tree, next to this file)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "adscf.hip")
LIB = os.path.join(HERE, "libadscf_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 arithmetic is kept off for the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # fuse a*b+c only inside one source expression: every instantiation of a pass
    # (single-instance, fixed-variant and batched kernels) then rounds identically
    "-ffp-contract=on",
]


def build(verbose: bool = False) -> str:
    """Compile the library; returns its path.  Raises CalledProcessError on failure."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, SRC, "-lhipfft"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise subprocess.CalledProcessError(res.returncode, cmd, output=out)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
