"""In-tree build of the native libraries (no cmake/ninja needed).

  lib/libdgr_bssn.so          HIP kernels + C-ABI + C++ shims, gfx950
  lib/libdgr_algebra_host.so  host build of the generated equations, used by
                              the CPU tests to check the generated code itself
  ../tests/cpp/shim_driver    test program calling the reference-signature
                              entry points (bssnRHS, bssnrhs) of libdgr_bssn.so

Usage: python build.py [--regen]   (--regen reruns the equation generator)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DGR_OFFLOAD_ARCH", "gfx950")
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def regenerate(verbose=False):
    _run([sys.executable, os.path.join(HERE, "codegen", "gen_bssn_rhs.py"),
          os.path.join(HERE, "csrc", "bssn_rhs_generated.h")], verbose)


# The FP64 kernels (and the rest of the library) are scheduled with the
# iterative ILP strategy; the FP32 fused kernels, compiled from the same
# source with DGR_F32_KERNELS in a unit of their own, keep the default one.
MAIN_FLAGS = ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]


def build_bssn(src: str, out: str, verbose: bool = False, extra=()) -> None:
    """libdgr_bssn.so from the HIP source `src` (two units) and the C++ shim."""
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]
    objs = [out + ".main.o", out + ".f32.o"]
    # the two units compile in parallel
    cmds = [[HIPCC, *HIP_FLAGS, *MAIN_FLAGS, *extra, "-c", *inc, src, "-o", objs[0]],
            [HIPCC, *HIP_FLAGS, *extra, "-DDGR_F32_KERNELS", "-c", *inc, src, "-o", objs[1]]]
    procs = []
    for c in cmds:
        if verbose:
            print(" ".join(c), flush=True)
        procs.append(subprocess.Popen(c))
    rcs = [p.wait() for p in procs]
    for c, rc in zip(cmds, rcs):
        if rc:
            raise subprocess.CalledProcessError(rc, c)
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", *inc, *objs, os.path.join(HERE, "host", "bssn_shim.cpp"), "-o", tmp],
         verbose)
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)


def build(verbose: bool = False) -> None:
    os.makedirs(LIB, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]  # for the host library
    build_bssn(os.path.join(HERE, "csrc", "dgr_bssn.hip"), os.path.join(LIB, "libdgr_bssn.so"), verbose)
    host = os.path.join(LIB, "libdgr_algebra_host.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", *inc,
          os.path.join(HERE, "csrc", "algebra_host.cpp"), "-o", host + ".tmp"], verbose)
    os.replace(host + ".tmp", host)
    build_shim_driver(verbose)


def build_shim_driver(verbose: bool = False) -> None:
    """tests/cpp/shim_driver: host program linked against lib/libdgr_bssn.so."""
    src = os.path.join(ROOT, "tests", "cpp", "shim_driver.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "shim_driver")
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-L", LIB, "-ldgr_bssn",
          "-Wl,-rpath,$ORIGIN/../../dendro-gr_amd/lib", "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)


if __name__ == "__main__":
    if "--regen" in sys.argv:
        regenerate(verbose=True)
    build(verbose=True)
