"""Builds libhpc_ode.so in-tree with hipcc for gfx950.

    python -m hpc_ode_amd.build            # compile everything
    python -m hpc_ode_amd.build --verbose

The kernels are compiled with packed FP32 disabled (the packed-fp32-ops target
feature off) and with machine-level loop-invariant code motion off: the latter
otherwise hoists the FP64 polynomial constants of exp/log out of the persistent
cell loop, where they stay live and push the BDF kernel into scratch spills.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.environ.get("HPO_CSRC") or os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.environ.get("HPO_LIB") or os.path.join(HERE, "libhpc_ode.so")
EXTRA = os.environ.get("HPO_EXTRA_FLAGS", "").split()
ARCH = os.environ.get("HPO_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-std=c++17", "-fPIC", "-I", CSRC, "-I", os.path.join(ROOT, "include")]
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-disable-machine-licm", "-x", "hip"]
HOST_FLAGS = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-O2", "-Wall", "-Wno-unused-result"]

SOURCES = [
    ("yaml_lite.cpp", HOST_FLAGS),
    ("mechanism.cpp", HOST_FLAGS),
    ("capi.cpp", HOST_FLAGS),
    ("integrator.cpp", HOST_FLAGS),
    ("bdf_kernel.hip", HIP_FLAGS),
]


def _compile(src, flags, verbose):
    obj = os.path.join(BUILD, os.path.basename(src) + (".x" if EXTRA else "") + ".o")
    cmd = [HIPCC, *COMMON, *flags, *EXTRA, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{out}")
    if verbose and out.strip():
        print(out)
    return obj


def build(verbose: bool = False) -> str:
    """Compiles every source and links libhpc_ode.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [(s, f) for s, f in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(lambda sf: _compile(sf[0], sf[1], verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("linking libhpc_ode.so failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build(verbose=a.verbose))


if __name__ == "__main__":
    sys.exit(main())
