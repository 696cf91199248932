"""Builds the native libraries in-tree.

- implicitcam_amd/_lib/libicam_hip.so: the product (HIP kernels for gfx950 and the
  C ABI of include/icam_hip.h), compiled by hipcc;
- implicitcam_amd/_lib/libicam_gen.so: the synthetic part generator (g++);
- implicitcam_amd/_lib/libicam_rules.so: the host build of the brick rules the device
  runs (csrc/brick_rules.hpp), which the tests check on the CPU (g++).

Every call compiles again (no cache), so a changed source is always picked up.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "_lib")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 instructions are left out of the device code (the compiler would
# otherwise form them from plain float code).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-munsafe-fp-atomics"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")
    return r


def build_gen(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libicam_gen.so")
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out + ".tmp", os.path.join(CSRC, "gen", "partgen.cpp")], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_rules(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libicam_rules.so")
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out + ".tmp", os.path.join(CSRC, "brick_rules.cpp")], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_hip(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libicam_hip.so")
    srcs = [os.path.join(CSRC, "icam_hip.hip"), os.path.join(CSRC, "cast.hip"), os.path.join(CSRC, "pack.cpp")]
    _run([HIPCC] + HIP_FLAGS + ["-o", out + ".tmp"] + srcs, verbose)
    os.replace(out + ".tmp", out)
    return out


def build_all(verbose=False):
    build_gen(verbose)
    build_rules(verbose)
    build_hip(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
