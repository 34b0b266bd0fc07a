"""Build libd2q9.so (gfx950) from pylabolt_amd/csrc with hipcc.

``python -m pylabolt_amd.build`` compiles every ``*.hip`` source into one
shared library next to this file.  Flags:

* ``-ffp-contract=off``: no fused multiply-add, so the FP64 path rounds
  every operation like the reference's scalar code does;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  instructions are generated.
"""
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libd2q9.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip"))


def _compile(src, extra, verbose):
    obj = os.path.join(CSRC, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], "-c", src,
           "-o", obj, *extra]
    t0 = time.time()
    subprocess.run(cmd, check=True)
    if verbose:
        print(f"[build] {os.path.basename(src)} {time.time() - t0:.1f}s",
              flush=True)
    return obj


def build(verbose=True, extra=()):
    """Compile all HIP sources (in parallel) into libd2q9.so; raises on
    failure."""
    t0 = time.time()
    srcs = sources()
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(lambda s: _compile(s, extra, verbose), srcs))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC",
                    *objs, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    if verbose:
        print(f"[build] {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=tuple(sys.argv[1:]))
