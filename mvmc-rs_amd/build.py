"""Build recipes (explicit hipcc / g++ command lines, in-tree outputs).

  libmvmc_hip.so   mvmc-rs_amd/_build/   HIP kernels for gfx950 + C ABI + host loader
  liboracle.so     oracle/_build/        CPU FP64 oracle (test infrastructure)

Floating-point contraction is disabled everywhere (-ffp-contract=off): every
parity path relies on unfused multiply/add sequences.  Packed FP32 code
generation is disabled on the device.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
ORACLE_BUILD = os.path.join(ROOT, "oracle", "_build")
LIB = os.path.join(BUILD, "libmvmc_hip.so")
ORACLE_LIB = os.path.join(ORACLE_BUILD, "liboracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MVMC_OFFLOAD_ARCH", "gfx950")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library(defines=(), out=None):
    """defines: extra -D macros (MVMC_PHASE_CLOCKS adds the sampling kernel's
    phase clocks); out: output path (default mvmc-rs_amd/_build/libmvmc_hip.so)."""
    out = out or LIB
    os.makedirs(os.path.dirname(out), exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", "capi.hip"), os.path.join(HERE, "host", "model.cpp")]
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-shared", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-result"] + [f"-D{d}" for d in defines] + ["-o", tmp] + srcs + \
          ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_oracle():
    os.makedirs(ORACLE_BUILD, exist_ok=True)
    srcs = [os.path.join(ROOT, "oracle", "oracle.cpp"), os.path.join(ROOT, "oracle", "expert.cpp")]
    tmp = ORACLE_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-o", tmp] + srcs + ["-lpthread"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


if __name__ == "__main__":
    what = sys.argv[1:] or ["oracle", "lib"]
    if "oracle" in what:
        build_oracle()
    if "lib" in what:
        build_library()
