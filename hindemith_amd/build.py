"""Compile libhmx.so (gfx950) in-tree with hipcc.

``python -m hindemith_amd.build`` or ``build()``: every ``csrc/*.hip`` is
compiled to an object under ``csrc/_build/`` (in parallel) and linked into
``hindemith_amd/libhmx.so``.  Always recompiles (the build check must prove the
sources compile); ``build(incremental=True)`` skips up-to-date objects.

Flags:
* ``--offload-arch=gfx950``: MI355X only, no other targets.
* ``-ffp-contract=off``: the kernels spell out every fused multiply-add
  (stencil taps), everything else rounds per op like the reference program.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libhmx.so")
REPO = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", os.path.join(REPO, "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src, incremental):
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + ".o")
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers.append(os.path.join(REPO, "include", "hmx.h"))
    if incremental and os.path.exists(obj):
        newest = max(os.path.getmtime(p) for p in [src] + headers)
        if os.path.getmtime(obj) >= newest:
            return obj, ""
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{' '.join(cmd)}\n"
                           f"{r.stdout}\n{r.stderr}")
    # hipcc's host pass warns that it does not know the packed-fp32 feature
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    return obj, warn


def build(incremental=False, verbose=False, jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, incremental) for s in srcs]
        for f in futs:
            obj, warn = f.result()
            objs.append(obj)
            if verbose and warn.strip():
                print(warn, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(incremental="--incremental" in sys.argv, verbose=True))
