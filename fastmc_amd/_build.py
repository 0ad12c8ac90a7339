"""Build libfastmc_hip.so in-tree with hipcc for gfx950."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfastmc_hip.so")
SOURCES = [os.path.join(CSRC, "fastmc_hip.hip")]
HEADERS = [os.path.join(CSRC, "fmc_device.hpp"), os.path.join(CSRC, "fmc_kernels.hpp"),
           os.path.join(CSRC, "fmc_grad.hpp"), os.path.join(CSRC, "fmc_widom.hpp"),
           os.path.join(os.path.dirname(HERE), "include", "fastmc.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FMC_OFFLOAD_ARCH", "gfx950")

FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
         # keep hand-written FP32 code off the packed-FP32 instructions
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


LIB_PHASES = os.path.join(HERE, "libfastmc_hip_phases.so")


def build(verbose: bool = False, phases: bool = False) -> str:
    """Compile every source (always) and return the library path.  phases:
    the profiling build with k_run's step clock (FMC_PHASE_CLOCK), written to
    libfastmc_hip_phases.so."""
    lib = LIB_PHASES if phases else LIB
    tmp = lib + ".tmp"
    cmd = [HIPCC, *FLAGS, *(["-DFMC_PHASE_CLOCK"] if phases else []), "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libfastmc_hip.so")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose=True, phases="--phases" in sys.argv))
