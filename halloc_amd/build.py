"""Build recipe of the MI355X halloc libraries (hipcc only, no cmake).

Produces, in-tree:
  halloc_amd/lib/libhalloc.a      relocatable-device-code archive with hamalloc /
                                  hafree / ha_init (what user code links with
                                  `-fgpu-rdc --hip-link -lhalloc`)
  halloc_amd/libhalloc_amd.so     the C-ABI library: libhalloc.a device-linked
                                  with the workload drivers (harness.hip)
Run `python -m halloc_amd.build` or call build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "build")
LIBDIR = os.path.join(PKG, "lib")
ARCHIVE = os.path.join(LIBDIR, "libhalloc.a")
SHARED = os.path.join(PKG, "libhalloc_amd.so")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
LLVM_AR = "/opt/rocm/lib/llvm/bin/llvm-ar"

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fgpu-rdc",
    # never let the compiler form packed fp32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines() if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build(verbose: bool = False) -> str:
    """Compile both libraries for gfx950; returns the path of libhalloc_amd.so."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    halloc_o = os.path.join(OBJ, "halloc.o")
    harness_o = os.path.join(OBJ, "harness.o")
    steps = [
        [HIPCC, *COMMON, "-c", os.path.join(CSRC, "halloc.hip"), "-o", halloc_o],
        [HIPCC, *COMMON, "-c", os.path.join(CSRC, "harness.hip"), "-o", harness_o],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
    if os.path.exists(ARCHIVE):
        os.remove(ARCHIVE)
    _run([LLVM_AR, "rcs", ARCHIVE, halloc_o])
    tmp = SHARED + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-fgpu-rdc", "--hip-link", "-shared", "-fPIC",
          harness_o, f"-L{LIBDIR}", "-lhalloc", "-o", tmp])
    shutil.move(tmp, SHARED)
    return SHARED


if __name__ == "__main__":
    print(build(verbose=True))
