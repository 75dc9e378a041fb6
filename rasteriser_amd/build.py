"""Build librz.so for gfx950, in-tree (rasteriser_amd/_lib/librz.so).

Arithmetic flags are part of correctness: -ffp-contract=off (no FMA contraction of
the reference's separate multiplies and adds), IEEE-correct f32 divide and sqrt, and
packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
LIB = os.path.join(OUT_DIR, "librz.so")
SOURCES = ["rz_kernels.hip", "rz_mlp.hip", "rz_api.hip", "rz_host_math.cpp"]
HOST_DIR = os.path.join(HERE, "host")
HOST_SOURCES = ["drawing.cpp", "fileloader.cpp", "image_io.cpp"]
HOST_LIB = os.path.join(OUT_DIR, "librz_host.so")
RENDERER = os.path.join(OUT_DIR, "renderer")
# The host layer is ordinary x86-64 C++ with the same arithmetic contract.
HOST_FLAGS = ["-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wextra"]
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIP_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops",)
    err = "\n".join(l for l in r.stderr.splitlines() if not any(n in l for n in noise))
    if r.returncode != 0:
        raise RuntimeError(f"librz build failed:\n{' '.join(cmd)}\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    build_host()
    return LIB


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed:\n{' '.join(cmd)}\n{r.stderr}")


def build_host() -> None:
    """librz_host.so (draw_frame, loaders, image I/O) and the renderer command line."""
    link = ["-L", OUT_DIR, "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    _run(["g++"] + HOST_FLAGS + ["-fPIC", "-shared", "-o", HOST_LIB + ".tmp"]
         + [os.path.join(HOST_DIR, s) for s in HOST_SOURCES] + link + ["-lrz", "-lz"])
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    _run(["g++"] + HOST_FLAGS + ["-o", RENDERER + ".tmp", os.path.join(HOST_DIR, "renderer.cpp")]
         + link + ["-lrz_host", "-lrz", "-lz"])
    os.replace(RENDERER + ".tmp", RENDERER)


if __name__ == "__main__":
    print(build(verbose=True))
