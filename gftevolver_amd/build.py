"""Builds the HIP library (libgft.so, gfx950) and, for the tests, the CPU checkers.

    python -m gftevolver_amd.build            # library + oracle (+ reference build when present)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgft.so")
REFERENCE = "/root/reference"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # Packed FP32 instructions stay off: the compiler must not emit them.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Contraction per source expression only (HIP's default also lets the
    # backend fuse across expressions, where the product it picks depends on
    # the surrounding code): the same expression rounds the same way in every
    # kernel variant, which the bitwise launch-shape and staging tests rely on.
    "-ffp-contract=on",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_library(verbose: bool = False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp, os.path.join(CSRC, "gft_capi.hip")]
    print("+", " ".join(cmd), flush=True)
    # The compiler's warnings (the host pass does not know the packed-FP32
    # target feature) stay quiet unless verbose; its errors are always shown.
    r = subprocess.run(cmd, stderr=None if verbose else subprocess.PIPE, text=True)
    if r.returncode != 0:
        if not verbose:
            sys.stderr.write(r.stderr)
        raise RuntimeError(f"hipcc failed ({r.returncode}) building {LIB}")
    os.replace(tmp, LIB)


def build_host():
    """C++ host API (gftevolver_amd/host) and the gft_run driver, linked to libgft.so."""
    host = os.path.join(HERE, "host")
    out = os.path.join(HERE, "gft_run")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", out + ".tmp", os.path.join(host, "gft_host.cpp"),
          os.path.join(host, "gft_run.cpp"), "-L" + HERE, "-lgft", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)


def build_checkers():
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", oracle, "oracle"])
    if os.path.isdir(REFERENCE):
        _run(["make", "-s", "-C", oracle, "ref", "-j8"])


def build_all():
    build_library()
    build_host()
    build_checkers()


if __name__ == "__main__":
    build_library(verbose="-v" in sys.argv)
    build_host()
    build_checkers()
