"""In-tree build of the solver.

  libdrot.so         hand-written gfx950 HIP kernels + C++ host driver + C ABI
                     (csrc/*.hip, include/drot.h), built with hipcc.
  reg_drot/_C.so     PyTorch-ROCm binding (reg_drot/_C.cpp) over the C ABI,
                     built with g++ against torch's headers.

Usage: python build.py [--lib-only]
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG_DIR)
INCLUDE = os.path.join(REPO, "include")
CSRC = os.path.join(PKG_DIR, "csrc")
LIB = os.path.join(PKG_DIR, "libdrot.so")
EXT = os.path.join(PKG_DIR, "reg_drot", "_C.so")

ARCH = os.environ.get("DROT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every fused multiply-add in the kernels is explicit, so
# results match the oracle's operation order.  Packed fp32 instructions are
# disabled for the device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-I", INCLUDE,
]

SOURCES = ["drot_kernels.hip", "drot_qr.hip", "drot_batch.hip", "drot_backward.hip",
           "drot_cost.hip", "drot_solver.hip"]


def _run(cmd, quiet_filter=()):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    lines = [ln for ln in proc.stdout.splitlines()
             if not any(f in ln for f in quiet_filter)]
    if proc.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def build_lib() -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs]
    _run(cmd, quiet_filter=("packed-fp32-ops' is not a recognized feature",))
    os.replace(tmp, LIB)
    return LIB


def build_ext() -> str:
    import torch
    from torch.utils import cpp_extension as ce

    incs = ce.include_paths() + [sysconfig.get_paths()["include"], INCLUDE,
                                 "/opt/rocm/include"]
    libdirs = ce.library_paths()
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
           "-DTORCH_EXTENSION_NAME=_C", "-DTORCH_API_INCLUDE_EXTENSION_H",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-w"]
    for i in incs:
        cmd += ["-I", i]
    cmd += [os.path.join(PKG_DIR, "reg_drot", "_C.cpp"), "-o", EXT + ".tmp"]
    for d in libdirs:
        cmd += ["-L", d, f"-Wl,-rpath,{d}"]
    cmd += ["-L", PKG_DIR, "-Wl,-rpath,$ORIGIN/..", "-ldrot",
            "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python"]
    _run(cmd)
    os.replace(EXT + ".tmp", EXT)
    return EXT


CLI = os.path.join(PKG_DIR, "drot_cli")


def build_cli() -> str:
    """The experiment command line (cli/drot_cli.cpp), a host program over libdrot.so."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE,
           os.path.join(PKG_DIR, "cli", "drot_cli.cpp"), "-o", CLI + ".tmp",
           "-L", PKG_DIR, "-Wl,-rpath,$ORIGIN", "-ldrot"]
    _run(cmd)
    os.replace(CLI + ".tmp", CLI)
    return CLI


def build_all(with_ext: bool = True) -> None:
    build_lib()
    build_cli()
    if with_ext:
        build_ext()


if __name__ == "__main__":
    build_all(with_ext="--lib-only" not in sys.argv)
    print("built", LIB, "" if "--lib-only" in sys.argv else EXT)
