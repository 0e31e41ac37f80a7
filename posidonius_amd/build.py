"""Build the in-tree HIP library ``libposidonius_hip.so`` for gfx950.

``python -m posidonius_amd.build`` (or ``__graft_entry__.build()``) compiles
``csrc/posidonius_hip.hip`` with hipcc.  Flags:

* ``-ffp-contract=off``: every multiply and add rounds separately, as in the
  reference (Rust never fuses), which is what makes bitwise parity with the
  CPU oracle possible;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed-FP32
  instructions are generated;
* ``--offload-arch=gfx950``: MI355X only;
* ``-mllvm -amdgpu-sched-strategy=iterative-maxocc``: the machine scheduler
  variant that measured fastest for the step kernel (3.15e8 vs 3.01e8
  system-steps/s with the default strategy on the 65,536-system ensemble).
  Scheduling reorders instructions only; results are bitwise unchanged.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "posidonius_hip.hip")
HEADERS = [os.path.join(HERE, "csrc", "whfast_step.hpp"),
           os.path.join(os.path.dirname(HERE), "include", "posidonius_hip.h")]
OUT = os.path.join(HERE, "libposidonius_hip.so")
ARCH = os.environ.get("POSIDONIUS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False, extra_flags=None) -> str:
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-mllvm", "-amdgpu-sched-strategy=iterative-maxocc",
           "-Wall", "-Wno-unused-function",
           "-I", os.path.join(os.path.dirname(HERE), "include"),
           SRC, "-o", OUT + ".tmp"]
    if extra_flags:
        cmd[1:1] = list(extra_flags)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


CLI_SRC = os.path.join(HERE, "host", "posidonius_cli.cpp")
CLI_OUT = os.path.join(HERE, "posidonius-hip")


def build_cli(verbose: bool = False) -> str:
    """The `posidonius-hip start|resume` host, linked against the library."""
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(hipcc())))
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(os.path.dirname(HERE), "include"), CLI_SRC,
           "-o", CLI_OUT + ".tmp", "-L", HERE, "-lposidonius_hip", "-Wl,-rpath,$ORIGIN",
           "-L", os.path.join(rocm, "lib"), "-Wl,-rpath," + os.path.join(rocm, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(CLI_OUT + ".tmp", CLI_OUT)
    return CLI_OUT


if __name__ == "__main__":
    print(build_library(verbose=True, extra_flags=sys.argv[1:]))
    print(build_cli(verbose=True))
