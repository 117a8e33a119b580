"""Build the native library ``openems_amd/lib/libopenems_hip.so`` for gfx950.

One shared library holds the host operator builder (include/ems_operator.h)
and the HIP engine (include/ems_hip.h).  Every source is compiled with
``-ffp-contract=off`` so that no multiply-add pair is fused (the engine is
bit-identical to the reference's separately rounded float arithmetic), and
with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libopenems_hip.so")
SOURCES = [
    os.path.join(HERE, "csrc", "host", "ems_operator.cpp"),
    os.path.join(HERE, "csrc", "engine", "ems_compress.cpp"),
    os.path.join(HERE, "csrc", "engine", "ems_engine.hip"),
]
ARCH = os.environ.get("OPENEMS_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def common_flags():
    return ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
            "-Wno-unused-result", "-pthread"]


def device_flags():
    return common_flags() + [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, os.path.basename(src) + ".o")
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-x", "c++"] + common_flags() + ["-c", src, "-o", obj]
        else:
            cmd = [HIPCC] + device_flags() + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, stderr=None if verbose else subprocess.PIPE)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-pthread", "-L/opt/rocm/lib",
                                                                          "-lrocprofiler-sdk-roctx"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_mirror_check(verbose)
    build_hbm_probe(verbose)
    return LIB


PROBE_SRC = os.path.join(HERE, "csrc", "bench", "ems_hbm_probe.hip")
PROBE_LIB = os.path.join(LIBDIR, "libems_hbm_probe.so")


def build_hbm_probe(verbose=False):
    """The HBM bandwidth probe bench.py reports beside the roofline (a library of
    its own: measurement, not part of the engine's ABI)."""
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC] + device_flags() + ["-shared", PROBE_SRC, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stderr=None if verbose else subprocess.PIPE)
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


MIRROR_SRC = os.path.join(ROOT, "tests", "cpp", "engine_mirror.cpp")
MIRROR_BIN = os.path.join(LIBDIR, "engine_mirror")


def build_mirror_check(verbose=False):
    """The C++ host-class check program (include/openems_hip.hpp), linked
    against the library; run by the GPU tests."""
    if not os.path.exists(MIRROR_SRC):
        return None
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), MIRROR_SRC,
           "-o", MIRROR_BIN, "-L", LIBDIR, "-lopenems_hip", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stderr=None if verbose else subprocess.PIPE)
    return MIRROR_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
