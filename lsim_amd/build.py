"""Build recipe for the native library lsim_amd/liblsim.so (gfx950).

hipcc compiles the HIP engine and the C++ host code (reader, levelizer,
generator, C ABI) into one shared library with the C ABI of include/lsim.h.
Objects go to lsim_amd/_build/ (not tracked); the .so is written in-tree so it
travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblsim.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["engine.hip", "netlist.cpp", "generator.cpp", "loops.cpp", "ingest.cpp", "producers.cpp", "capi.cpp"]
# Measurement-only library (random-read probe) loaded by bench.py.
PROBE_SOURCES = ["probe.hip"]
PROBE_LIB = os.path.join(HERE, "liblsim_probe.so")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]
# Device code: gfx950 only; packed-FP32 instruction forms disabled.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", path, "-o", out]
    else:
        cmd = [HIPCC] + COMMON + ["-x", "c++", "-c", path, "-o", out]
    subprocess.run(cmd, check=True)
    return out


def build_facade_check() -> str:
    """The C++ facade test driver (tests/cpp/facade_check.cpp), linked to liblsim.so."""
    src = os.path.join(ROOT, "tests", "cpp", "facade_check.cpp")
    out_dir = os.path.join(ROOT, "tests", "cpp", "_bin")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "facade_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), "-o", out, src,
                    "-L" + HERE, "-llsim", "-Wl,-rpath,$ORIGIN/../../../lsim_amd"], check=True)
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES) + len(PROBE_SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES + PROBE_SOURCES))
    for lib_path, lib_objs in ((LIB, objs[:len(SOURCES)]), (PROBE_LIB, objs[len(SOURCES):])):
        tmp = lib_path + ".tmp"
        cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + lib_objs + ["-lpthread"]
        subprocess.run(cmd, check=True)
        os.replace(tmp, lib_path)
    build_facade_check()
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
