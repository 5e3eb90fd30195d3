"""In-tree build of the native libraries.

    python -m minimumcyclebasis_amd.build [--no-oracle]

* ``lib/libmcb_host.so``  g++ -O2 -fopenmp (host preprocessing, include/mcb_host.h)
* ``lib/libmcb_hip.so``   hipcc --offload-arch=gfx950 (kernels + device C ABI, include/mcb_hip.h)
* ``oracle/liboracle.so`` and, when the reference tree is present, ``oracle/_ref/*`` (the checker;
  built here so that it travels with the tree, never linked into the product).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "lib")
HOST_SRC = [os.path.join(PKG, "csrc", "host", "mcb_host.cpp")]
HIP_SRC = [os.path.join(PKG, "csrc", "hip", "mcb_hip.hip"), os.path.join(PKG, "csrc", "hip", "mcb_mlp.hip"),
           os.path.join(PKG, "csrc", "hip", "mcb_sptree.hip")]
ARCH = os.environ.get("MCB_OFFLOAD_ARCH", "gfx950")

HOST_FLAGS = ["-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{INC}"]
# Packed FP32 instructions stay disabled in every HIP object.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-Wall", f"-I{INC}"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_host() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmcb_host.so")
    tmp = out + ".tmp"
    _run(["g++", *HOST_FLAGS, "-o", tmp, *HOST_SRC])
    os.replace(tmp, out)
    return out


def build_hip() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmcb_hip.so")
    tmp = out + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, "-o", tmp, *HIP_SRC])
    os.replace(tmp, out)
    return out


def build_oracle() -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"])


def build_all(oracle: bool = True) -> None:
    build_host()
    build_hip()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build_all(oracle="--no-oracle" not in sys.argv)
