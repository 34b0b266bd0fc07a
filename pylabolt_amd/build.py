"""Build libd2q9.so (gfx950) from pylabolt_amd/csrc with hipcc.

``python -m pylabolt_amd.build`` compiles every ``*.hip`` source into one
shared library next to this file.  Flags:

* ``-ffp-contract=off``: no fused multiply-add, so the FP64 path rounds
  every operation like the reference's scalar code does;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  instructions are generated.

The library carries a digest of the sources and flags it was built from
(``d2q9_source_hash()``); ``_native.load_library`` refuses a library whose
digest differs from that of the tree it is loaded from.
"""
import hashlib
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libd2q9.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "d2q9.h")
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


def source_digest(csrc=CSRC, header=HEADER):
    """Digest of what a build compiles: the compiler flags, every ``.hip``
    and ``.h`` file of ``csrc`` (by name and content) and the C header."""
    h = hashlib.sha256()
    h.update(" ".join(FLAGS).encode())
    names = sorted(f for f in os.listdir(csrc) if f.endswith((".hip", ".h")))
    files = [(n, os.path.join(csrc, n)) for n in names]
    files.append(("include/d2q9.h", header))
    for name, path in files:
        with open(path, "rb") as fh:
            data = fh.read()
        h.update(b"\0" + name.encode() + b"\0" + data)
    return h.hexdigest()[:32]


def _object_key(src, extra):
    """Digest of one object's inputs: its source, every header of csrc and
    the C header (any source may include them), the flags."""
    h = hashlib.sha256()
    h.update(" ".join(FLAGS + list(extra)).encode())
    names = sorted(f for f in os.listdir(CSRC) if f.endswith(".h"))
    for path in [src] + [os.path.join(CSRC, n) for n in names] + [HEADER]:
        with open(path, "rb") as fh:
            h.update(b"\0" + os.path.basename(path).encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def _compile(src, extra, verbose, digest, obj_dir):
    base = os.path.basename(src)[:-4]
    stamped = base == "d2q9_version"
    if obj_dir is None or stamped:
        obj = os.path.join(CSRC, base + ".o")
    else:
        obj = os.path.join(obj_dir, f"{base}.{_object_key(src, extra)}.o")
        if os.path.exists(obj):
            return obj, False
    cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], "-c", src,
           "-o", obj, *extra]
    if stamped:
        cmd.insert(1, f"-DD2Q9_SOURCE_HASH=h{digest}")
    t0 = time.time()
    subprocess.run(cmd, check=True)
    if verbose:
        print(f"[build] {os.path.basename(src)} {time.time() - t0:.1f}s",
              flush=True)
    return obj, obj_dir is None or stamped


def build(verbose=True, extra=(), reuse_objects=False):
    """Compile all HIP sources (in parallel) into libd2q9.so; raises on
    failure.  With ``reuse_objects`` (a development aid; the default
    compiles everything), objects whose source, headers and flags are
    unchanged are taken from ``csrc/_obj``."""
    t0 = time.time()
    srcs = sources()
    digest = source_digest()
    obj_dir = None
    if reuse_objects:
        obj_dir = os.path.join(CSRC, "_obj")
        os.makedirs(obj_dir, exist_ok=True)
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(jobs) as pool:
        res = list(pool.map(
            lambda s: _compile(s, extra, verbose, digest, obj_dir), srcs))
    objs = [o for o, _ in res]
    if obj_dir is not None:
        # keep only the objects of the current sources
        keep = {os.path.basename(o) for o in objs}
        for f in os.listdir(obj_dir):
            if f.endswith(".o") and f not in keep:
                os.remove(os.path.join(obj_dir, f))
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC",
                    *objs, "-o", tmp], check=True)
    os.replace(tmp, LIB)
    for obj, temporary in res:
        if temporary:
            os.remove(obj)
    if verbose:
        print(f"[build] {LIB} in {time.time() - t0:.1f}s", flush=True)
    return LIB


if __name__ == "__main__":
    args = sys.argv[1:]
    reuse = "--reuse" in args
    build(verbose=True, extra=tuple(a for a in args if a != "--reuse"),
          reuse_objects=reuse)
