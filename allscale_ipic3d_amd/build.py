"""Builds the gfx950 HIP libraries of the package in-tree (allscale_ipic3d_amd/_lib/).

    python -m allscale_ipic3d_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "_lib")
ARCH = os.environ.get("IPIC_OFFLOAD_ARCH", "gfx950")

# packed FP32 instructions are disabled for all device code
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]

LIBS = {
    "libipic_gpu.so": ["csrc/ipic_gpu.hip", "csrc/ipic_nop.hip", "csrc/ipic_refinit.cpp",
                       "csrc/ipic_fieldinit.cpp"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    return "hipcc"


def build(verbose: bool = False) -> list[str]:
    os.makedirs(LIBDIR, exist_ok=True)
    out = []
    for lib, srcs in LIBS.items():
        dst = os.path.join(LIBDIR, lib)
        tmp = dst + ".tmp"
        cmd = [hipcc(), *HIPFLAGS, "-o", tmp, *[os.path.join(HERE, s) for s in srcs]]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed for {lib}")
        os.replace(tmp, dst)
        out.append(dst)
    return out


CPP_TESTS = {
    # C++ mirror of the reference's step API, driven by tests/test_cpp_mirror.py
    os.path.join(ROOT, "tests", "cpp", "ipic3d_mirror_test"): os.path.join(ROOT, "tests", "cpp", "ipic3d_mirror_test.cpp"),
}


def build_cpp(verbose: bool = False) -> list[str]:
    """Host programs over the C ABI (g++, linked against the in-tree libipic_gpu.so)."""
    out = []
    for dst, src in CPP_TESTS.items():
        rpath = os.path.relpath(LIBDIR, os.path.dirname(dst))
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", dst + ".tmp", src,
               "-L", LIBDIR, "-lipic_gpu", f"-Wl,-rpath,$ORIGIN/{rpath}"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed for {src}")
        os.replace(dst + ".tmp", dst)
        out.append(dst)
    return out


if __name__ == "__main__":
    for p in build(verbose=True) + build_cpp(verbose=True):
        print("built", p)
