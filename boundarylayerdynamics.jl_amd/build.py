"""Builds libbld.so (HIP kernels for gfx950 + the C ABI) in-tree with hipcc.

Every source (and every transform part of kernels.hip, selected with
BLD_KERNEL_PART) is compiled to an object of its own in parallel, then the
objects are linked into the shared library."""

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
KERNEL_PARTS = 17  # kernels.hip: part 0 (physics, stages, launchers) and one part per transform kind and type
UNITS = ([("csrc/kernels.hip", [f"-DBLD_KERNEL_PART={p}"]) for p in range(KERNEL_PARTS)]
         + [(s, []) for s in ("csrc/engine.hip", "csrc/mlp.hip", "csrc/closure.hip", "csrc/profiles.hip",
                              "csrc/noalloc.hip")])
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"),
]


def _jobs():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 4
    return max(1, min(n, 16, int(os.environ.get("MAX_JOBS", "16"))))


def build(verbose=False):
    out = os.path.join(HERE, "libbld.so")
    tmp = out + ".tmp"
    with tempfile.TemporaryDirectory(prefix="bld_build_") as objdir:
        def compile_unit(i):
            src, defs = UNITS[i]
            obj = os.path.join(objdir, f"u{i}.o")
            cmd = [HIPCC, *FLAGS, *defs, "-c", os.path.join(HERE, src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            return cmd, obj, r

        with ThreadPoolExecutor(_jobs()) as ex:
            results = list(ex.map(compile_unit, range(len(UNITS))))
        for cmd, _, r in results:
            if r.returncode != 0:
                sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
                raise RuntimeError("hipcc failed building libbld.so")
            if verbose:
                sys.stderr.write(" ".join(cmd) + "\n")
        link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *[o for _, o, _ in results], "-o", tmp]
        r = subprocess.run(link, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(" ".join(link) + "\n" + r.stdout + r.stderr)
            raise RuntimeError("hipcc failed linking libbld.so")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
