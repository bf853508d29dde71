"""Build the gfx950 HIP library and the ando_barrier_core module in-tree.

    python -m andosim_amd.build          # both
    python -m andosim_amd.build --lib    # only libandosim_hip.so

Products (next to this file):
  libandosim_hip.so                  C ABI of include/andosim_hip.h (hipcc, gfx950)
  ando_barrier_core<EXT_SUFFIX>      Python module mirroring the reference binding,
                                     linked against libandosim_hip.so (rpath $ORIGIN)
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libandosim_hip.so")
EXT = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
MODULE = os.path.join(HERE, "ando_barrier_core" + EXT)

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", os.path.join(ROCM, "bin", "hipcc"))
ARCH = os.environ.get("ANDO_OFFLOAD_ARCH", "gfx950")

# Rounding semantics of the reference's x86-64 float build: no contraction,
# correctly rounded fp32 divide and sqrt.  Packed-FP32 instructions are kept
# out of the generated code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    src = os.path.join(CSRC, "ctx.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-I", os.path.join(ROOT, "include"), "-o", tmp, src])
    os.replace(tmp, LIB)
    return LIB


def build_copy_probe():
    """Counter-calibration program (plain copy of known size); measurement only."""
    out = os.path.join(HERE, "copy_probe")
    _run([HIPCC, *HIP_FLAGS, "-o", out + ".tmp", os.path.join(CSRC, "copy_probe.hip")])
    os.replace(out + ".tmp", out)
    return out


def build_module():
    import pybind11

    src = os.path.join(CSRC, "bindings.cpp")
    tmp = MODULE + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-fvisibility=hidden",
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"], "-I", os.path.join(ROOT, "include"),
           "-I", CSRC, "-o", tmp, src, "-L", HERE, "-landosim_hip", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(tmp, MODULE)
    return MODULE


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--module", action="store_true")
    a = ap.parse_args(argv)
    both = not (a.lib or a.module)
    if a.lib or both:
        build_lib()
        build_copy_probe()
    if a.module or both:
        build_module()


if __name__ == "__main__":
    sys.exit(main())
