"""Build libsimpeg_dc.so (HIP, gfx950) in-tree.

``python -m simpeg_amd.build`` compiles ``simpeg_amd/csrc/*.hip`` with hipcc into
``simpeg_amd/libsimpeg_dc.so``.  Packed-fp32 instruction selection is disabled
(``-packed-fp32-ops``) for every HIP source.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libsimpeg_dc.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
ARCH = os.environ.get("SDC_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cand):
        raise RuntimeError("hipcc not found (ROCm is required to build libsimpeg_dc.so)")
    return cand


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    srcs = sources()
    tmp = LIB_PATH + ".tmp"
    cmd = [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result",
        "-I", INCLUDE,
        "-I", CSRC,
        "-o", tmp,
        *srcs,
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
