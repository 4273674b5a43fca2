"""Builds libmwdg2.so (HIP kernels for gfx950 + host C++ + C ABI) in-tree.

    python gpu-mwdg2_amd/build.py [--check]

The library lands next to this file; the command-line driver `mwdg2_run`
(the drop-in for the reference executable) is linked against it.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmwdg2.so")
EXE = os.path.join(HERE, "mwdg2_run")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: every multiply and add rounds on its own, on host and
# device, as the oracle does. Packed FP32 instructions are disabled.
COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    return [os.path.join(CSRC, f) for f in ("mw_capi.hip", "mw_input.cpp")]


def build(verbose: bool = False) -> str:
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in sources():
        obj = os.path.join(HERE, "_build", os.path.basename(src) + ".o")
        cmd = [HIPCC] + COMMON + (DEVICE + ["-x", "hip"] if src.endswith(".hip") else []) + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", tmp] + objs, check=True)
    os.replace(tmp, LIB)
    drv = os.path.join(CSRC, "mw_run.cpp")
    subprocess.run([HIPCC, "-std=c++17", "-O2", drv, "-o", EXE, "-L" + HERE, "-lmwdg2",
                    "-Wl,-rpath,$ORIGIN"], check=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
