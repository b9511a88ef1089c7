"""Builds liblmff_hip.so (HIP kernels + C ABI) for gfx950 in-tree with hipcc.

Usage: python lammps-lmff_amd/build.py [--force]
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(HERE, "liblmff_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["lmff_engine.hip", "potentials.cpp", "pair_lmff_hip.cpp"]
HEADERS = ["lmff_kernels.hpp", "potentials.hpp", "pair_lmff_hip.hpp"]
RUNNER = os.path.join(HERE, "lmff_run")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in any kernel
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}"]


def _stale() -> bool:
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    deps = [os.path.join(CSRC, f) for f in SOURCES + HEADERS + ["lmff_run.cpp"]] + [os.path.join(INCLUDE, "lmff_hip.h")]
    return any(os.path.getmtime(d) > t for d in deps)


def build(force: bool = True, verbose: bool = False) -> str:
    """Compiles every source; always recompiles when force is set."""
    if not force and not _stale():
        return LIB
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src + ".o")
        cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    # deck-subset runner (host C++ over the C ABI)
    cmd = [HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}", os.path.join(CSRC, "lmff_run.cpp"),
           f"-L{HERE}", "-llmff_hip", "-Wl,-rpath,$ORIGIN", "-o", RUNNER]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv or True, verbose=True)
    print(LIB)
