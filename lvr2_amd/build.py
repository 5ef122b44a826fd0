"""Build recipes for the native parts of lvr2_amd.

- liblvrhip.so   : the gfx950 HIP library behind include/lvrhip.h (hipcc)
- libscenes.so   : the seeded synthetic scene generators (g++/OpenMP, host only)

Run ``python -m lvr2_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

LIB = os.path.join(HERE, "liblvrhip.so")
SCENES = os.path.join(HERE, "libscenes.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: the kNN key d^2 and the ordered fp32 sums must round as written.
# Packed fp32 ops are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    src = os.path.join(CSRC, "lvrhip.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", src, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_scenes() -> str:
    src = os.path.join(CSRC, "scenes.cpp")
    tmp = SCENES + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", src, "-o", tmp])
    os.replace(tmp, SCENES)
    return SCENES


def build_all() -> None:
    build_scenes()
    build_lib()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "scenes": build_scenes}[w]()
