"""Build recipe of the HIP library (gfx950).

    python -m sister_amd.build          # -> sister_amd/libsister_hip.so

hipcc compiles the device kernels and the host runtime into one shared
library exporting the C ABI of include/sister.h.  All HIP code is built with
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsister_hip.so")
SOURCES = ["kernels.hip", "runtime.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SISTER_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose=False):
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(HERE, "_build", src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l]
        if r.returncode != 0:
            sys.stderr.write("\n".join(msgs) + "\n")
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose and msgs:
            print("\n".join(msgs))
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_cli(verbose)
    build_probe(verbose)
    return LIB


PROBE = os.path.join(HERE, "libvalu_peak.so")


def build_probe(verbose=False):
    """The VALU throughput probe (csrc/valu_peak.hip) as a library of its own:
    bench.py --valu-peak loads it; the product library does not contain it."""
    tmp = PROBE + ".tmp"
    cmd = [HIPCC, *FLAGS, "-shared", os.path.join(CSRC, "valu_peak.hip"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed on valu_peak.hip")
    os.replace(tmp, PROBE)
    return PROBE


CLI = os.path.join(HERE, "compute_disp")


def build_cli(verbose=False):
    """compute_disp: host C++ over the C ABI (g++, zlib)."""
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "cli", "compute_disp.cpp"),
           "-I", os.path.join(ROOT, "include"), "-L", HERE, "-lsister_hip", "-lz",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", CLI]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return CLI


if __name__ == "__main__":
    print(build(verbose=True))
