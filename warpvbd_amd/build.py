"""Compile the HIP sources under csrc/ into libwarpvbd.so for gfx950.

Run as ``python -m warpvbd_amd.build`` or through ``__graft_entry__.build()``.
Objects go to ``warpvbd_amd/_build/``; the shared library lands next to this
file, where ``warpvbd_amd._lib`` loads it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libwarpvbd.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
# -ffp-contract=on: a multiply and an add are fused only within one source
# expression, so the same arithmetic rounds the same way wherever it is inlined
# (with fusion across statements the staged and unstaged sweeps of one vertex
# differed in the last bit).
FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall", "-Wno-unused-function",
         "-ffp-contract=on",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(HERE, "..", "include")]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str, defines=(), build_dir=None) -> str:
    bdir = build_dir or BUILD
    obj = os.path.join(bdir, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False, defines=(), out: str = None) -> str:
    """Compile every csrc/*.hip (optionally with extra -D defines) and link
    `out` (default: warpvbd_amd/libwarpvbd.so)."""
    out = out or OUT
    bdir = BUILD if out == OUT else out + ".build"
    os.makedirs(bdir, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, defines, bdir), srcs))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
