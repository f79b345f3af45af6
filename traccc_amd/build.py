"""In-tree build of the traccc_amd libraries (no CMake).

  traccc_amd/_lib/libtraccc_amd.so      HIP kernels for gfx950 + the C ABI
  traccc_amd/_lib/libtraccc_amd_sim.so  host-only synthetic-input toolkit
  oracle/liboracle_ckf.so               CPU oracle (test infrastructure)

Run as `python -m traccc_amd.build [--no-hip] [--no-oracle]`.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "_lib")
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TRACCC_AMD_ARCH", "gfx950")

HIP_SOURCES = [
    os.path.join(CSRC, "ckf", "runtime.hip"),
    os.path.join(CSRC, "ckf", "kernels.hip"),
    os.path.join(CSRC, "ckf", "votes.hip"),
    os.path.join(CSRC, "ckf", "pkf.hip"),
    os.path.join(CSRC, "scorer", "mlp.hip"),
    os.path.join(CSRC, "fit", "fit.hip"),
    os.path.join(CSRC, "ambiguity", "greedy.hip"),
    os.path.join(CSRC, "seeding", "seeding.hip"),
    os.path.join(CSRC, "common", "config.cpp"),
    os.path.join(CSRC, "common", "last_error.cpp"),
    os.path.join(CSRC, "io", "csv.cpp"),
]
SIM_SOURCES = [
    os.path.join(CSRC, "sim", "simulation.cpp"),
    os.path.join(CSRC, "common", "config.cpp"),
]

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_sim() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libtraccc_amd_sim.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-Wno-missing-field-initializers", f"-I{INC}", *SIM_SOURCES, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_hip() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libtraccc_amd.so")
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(LIB_DIR, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC, *HIP_FLAGS, f"-I{INC}", f"-I{CSRC}", *lang, "-c", src, "-o", obj])
        objs.append(obj)
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    for o in objs:
        os.remove(o)
    return out


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle_ckf.so")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-hip", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)
    build_sim()
    if not a.no_oracle and shutil.which("make"):
        build_oracle()
    if not a.no_hip:
        build_hip()
    return 0


if __name__ == "__main__":
    sys.exit(main())
