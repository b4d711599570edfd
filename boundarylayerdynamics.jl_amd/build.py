"""Builds libbld.so (HIP kernels for gfx950 + the C ABI) in-tree with hipcc."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
SOURCES = ["csrc/kernels.hip", "csrc/engine.hip", "csrc/mlp.hip", "csrc/closure.hip", "csrc/profiles.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"),
]


def build(verbose=False):
    out = os.path.join(HERE, "libbld.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *[os.path.join(HERE, s) for s in SOURCES], "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libbld.so")
    if verbose:
        sys.stderr.write(" ".join(cmd) + "\n")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
