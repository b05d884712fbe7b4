"""Builds libklt.so (HIP for gfx950 + host C++) and the klitho_* tools in-tree.

    python k-litho_amd/build.py            # library + tools
Outputs: k-litho_amd/lib/libklt.so, k-litho_amd/bin/klitho_{tcc,socs}.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "lib", "obj")
LIB = os.path.join(PKG, "lib", "libklt.so")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every HIP translation unit.
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _sources():
    out = []
    for d in (CSRC,):
        for f in sorted(os.listdir(d)):
            if f.endswith((".hip", ".cpp")):
                out.append(os.path.join(d, f))
    return out


# Host translation units that repeat the reference's float decisions exactly.
NO_CONTRACT = {"host_io.cpp", "png_io.cpp", "tcc.hip", "abbe.hip", "resist.hip"}


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    extra = ["-ffp-contract=off"] if os.path.basename(src) in NO_CONTRACT else []
    cmd = [HIPCC, *HIPFLAGS, *extra, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-x")
        cmd.insert(2, "hip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True, capture_output=not verbose)
    os.replace(tmp, LIB)
    tools = []
    if os.path.isdir(HOST):
        for tool in ("klitho_tcc", "klitho_socs"):
            src = os.path.join(HOST, tool + ".cpp")
            if not os.path.exists(src):
                continue
            exe = os.path.join(BIN, tool)
            common = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST))
                      if f.endswith(".cpp") and not f.startswith("klitho_")]
            cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
                   "-o", exe, src, *common, "-L" + os.path.dirname(LIB), "-lklt",
                   "-Wl,-rpath,$ORIGIN/../lib"]
            subprocess.run(cmd, check=True)
            tools.append(exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
