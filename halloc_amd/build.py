"""Build recipe of the MI355X halloc libraries (hipcc only, no cmake).

Produces, in-tree:
  halloc_amd/lib/libhalloc.a      relocatable-device-code archive with hamalloc /
                                  hafree / ha_init (what user code links with
                                  `-fgpu-rdc --hip-link -lhalloc`); the entry
                                  points inline into the calling kernels
  halloc_amd/lib/libhalloc_ool.a  the same with out-of-line entry points
                                  (-DHALLOC_OUT_OF_LINE), as the reference's
                                  __noinline__ hamalloc / hafree: one copy of
                                  the allocator for programs with many call sites
  halloc_amd/libhalloc_amd.so     the C-ABI library: libhalloc.a device-linked
                                  with the workload drivers (harness.hip)
  halloc_amd/libhalloc_amd_ool.so the same drivers linked against libhalloc_ool.a
                                  (select it with HALLOC_AMD_LIB to measure the
                                  out-of-line calls)
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
ARCHIVE_OOL = os.path.join(LIBDIR, "libhalloc_ool.a")
SHARED = os.path.join(PKG, "libhalloc_amd.so")
SHARED_OOL = os.path.join(PKG, "libhalloc_amd_ool.so")

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


def _link(harness_o: str, lib: str, out: str) -> None:
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-fgpu-rdc", "--hip-link", "-shared", "-fPIC",
          harness_o, f"-L{LIBDIR}", f"-l{lib}", "-o", tmp])
    shutil.move(tmp, out)


def build(verbose: bool = False) -> str:
    """Compile the libraries for gfx950; returns the path of libhalloc_amd.so."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    halloc_o = os.path.join(OBJ, "halloc.o")
    halloc_ool_o = os.path.join(OBJ, "halloc_ool.o")
    harness_o = os.path.join(OBJ, "harness.o")
    steps = [
        [HIPCC, *COMMON, "-c", os.path.join(CSRC, "halloc.hip"), "-o", halloc_o],
        [HIPCC, *COMMON, "-DHALLOC_OUT_OF_LINE", "-c", os.path.join(CSRC, "halloc.hip"), "-o", halloc_ool_o],
        [HIPCC, *COMMON, "-c", os.path.join(CSRC, "harness.hip"), "-o", harness_o],
    ]
    # the three compiles are independent: run them side by side
    procs = []
    for cmd in steps:
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for cmd, p in procs:
        out = "\n".join(line for line in p.communicate()[0].splitlines() if "packed-fp32-ops" not in line)
        if p.returncode != 0:
            raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}\n{out}")
        if out.strip():
            print(out, file=sys.stderr)
    for archive, obj in ((ARCHIVE, halloc_o), (ARCHIVE_OOL, halloc_ool_o)):
        if os.path.exists(archive):
            os.remove(archive)
        _run([LLVM_AR, "rcs", archive, obj])
    _link(harness_o, "halloc", SHARED)
    _link(harness_o, "halloc_ool", SHARED_OOL)
    return SHARED


if __name__ == "__main__":
    print(build(verbose=True))
