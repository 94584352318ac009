"""Builds the HIP library gravity_amd/libgravity_hip.so for gfx950 with hipcc.

Every source compiles with -ffp-contract=off (float expressions round as
written, as the CPU oracle's do) and with packed FP32 instructions disabled.
Objects go to gravity_amd/_build/; the library lands next to this file so it
travels with the tree.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
BUILD = PACKAGE / "_build"
LIBRARY = PACKAGE / "libgravity_hip.so"
SOURCES = ["sort.hip", "collide.hip", "p3m.hip", "fft.hip", "liquid.hip", "sleep.hip", "adjoint.hip", "trajectory.hip", "rigid.hip",
           "engine.hip", "green.cpp", "members.cpp", "tuning.cpp", "capi.cpp"]
ARCH = os.environ.get("GRAVITY_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return path


def _compile(source: str) -> Path:
    src = CSRC / source
    obj = BUILD / (source + ".o")
    lang = ["-x", "hip"] if source.endswith(".hip") else []
    cmd = [hipcc(), *lang, f"--offload-arch={ARCH}", *COMMON_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        objects = list(pool.map(_compile, SOURCES))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    try:
        build(verbose=True)
    except RuntimeError as err:
        print(err, file=sys.stderr)
        sys.exit(1)
