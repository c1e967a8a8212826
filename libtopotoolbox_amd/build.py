"""Build the HIP library in-tree: libtopotoolbox_amd/_lib/libtopotoolbox_amd.so.

One hipcc invocation compiles every kernel source for gfx950 and links the
shared library that carries the C ABI (include/topotoolbox_amd.h).
Run as ``python -m libtopotoolbox_amd.build`` or through
``__graft_entry__.build()``.
"""
from __future__ import annotations

import glob
import hashlib
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_lib")
LIB_NAME = "libtopotoolbox_amd.so"
LIB_PATH = os.path.join(OUT_DIR, LIB_NAME)

SOURCES = ["runtime.hip", "fill.hip", "fsm.hip", "fmm3d.hip", "adjoint.hip", "batched.hip",
           "stencil.hip", "gwdt.hip", "capi.hip"]
HEADERS = ["tile_engine.h", "kernels.h", "workspace.h"]

ARCH = os.environ.get("TT_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # every proposal is evaluated with the reference's float operations,
    # one rounding per operation: no fused multiply-add contraction
    "-ffp-contract=off",
    # keep packed FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def source_hash() -> str:
    """SHA-256 (first 32 hex digits) over the library's sources: every
    csrc/*.hip and csrc/*.h and the public header. Stamped into the built
    library (tt_build_hash) so a loader can refuse a stale binary."""
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.h")) +
                   glob.glob(os.path.join(ROOT, "include", "*.h")))
    for path in files:
        h.update(os.path.relpath(path, ROOT).encode())
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:32]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def sources() -> list:
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def _run(cmd: list, verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}: {' '.join(cmd)}")


def build(verbose: bool = False) -> str:
    """Compile the library (always) and return its path. Each source is
    compiled to an object in parallel (every translation unit carries its
    own device code), then the objects are linked into the shared library."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OUT_DIR, exist_ok=True)
    obj_dir = os.path.join(OUT_DIR, "obj")
    os.makedirs(obj_dir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    compile_flags = [f for f in FLAGS if f != "-shared"] + [f'-DTT_SOURCE_HASH="{source_hash()}"']
    jobs = []
    for src in sources():
        obj = os.path.join(obj_dir, os.path.basename(src) + ".o")
        jobs.append((obj, [hipcc(), *compile_flags, *inc, "-c", src, "-o", obj]))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, cmd, verbose) for _, cmd in jobs]:
            f.result()
    tmp = LIB_PATH + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *[o for o, _ in jobs], "-o",
          tmp], verbose)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
