"""Builds libsolar_amd.so (HIP kernels for gfx950 + host C-ABI) in-tree.

    python -m solar_amd.build        # or: python solar_amd/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsolar_amd.so")
ARCH = os.environ.get("SOLAR_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps every multiply and add separately rounded, as in the
# reference; packed FP32 instructions are disabled for the whole build.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I/opt/rocm/include"]
SOURCES_HIP = ["solar_kernels.hip", "solar_adjoint.hip", "solar_noalloc.hip"]
SOURCES_HOST = ["solar_tables.cpp", "solar_ic.cpp", "solar_capi.cpp", "solar_ensemble.cpp", "solar_io.cpp"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return "hipcc"


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES_HIP:
        obj = os.path.join(objdir, src + ".o")
        cmd = [hipcc(), *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd, verbose)
        objs.append(obj)
    for src in SOURCES_HOST:
        obj = os.path.join(objdir, src + ".o")
        # host-only translation units: no device code, no offload target
        cmd = [hipcc(), *HOST_FLAGS, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-c",
               os.path.join(CSRC, src), "-o", obj]
        _run(cmd, verbose)
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, OUT)
    return OUT


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    if verbose and (r.stdout or r.stderr):
        sys.stderr.write(r.stdout + r.stderr)


if __name__ == "__main__":
    print(build(verbose=True))
