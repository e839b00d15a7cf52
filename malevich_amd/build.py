"""Build the gfx950 shared library (the C ABI of include/malevich.h).

Compiles every HIP source under malevich_amd/csrc with hipcc for gfx950 and
links libmalevich_hip.so in-tree, next to this file, so it travels with the
repository snapshot. FP contraction is disabled: the pipeline reproduces the
reference's separate multiply/add rounding in edge, depth and attribute math.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libmalevich_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MLV_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]

SOURCES = ["mlv_kernels.hip", "mlv_runtime.hip", "mlv_diff.hip", "mlv_camera.hip"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")
    noise = "is not a recognized feature for this target"
    rest = "\n".join(l for l in (proc.stdout + proc.stderr).splitlines() if noise not in l and l.strip())
    if rest:
        sys.stderr.write(rest + "\n")


def build_library(verbose: bool = False) -> Path:
    """Compile and link libmalevich_hip.so (always rebuilds)."""
    BUILD.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd))
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-lrocprofiler-sdk-roctx", "-o",
          str(tmp)])
    os.replace(tmp, LIB)
    return LIB


HOST_PROGRAM = PKG / "mlv_render"


def build_host_program() -> Path:
    """Compile the C host program (malevich_amd/host/mlv_render.c) with gcc against the C ABI."""
    _run(["gcc", "-std=c11", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", str(PKG / "host" / "mlv_render.c"),
          f"-L{PKG}", "-lmalevich_hip", "-Wl,-rpath,$ORIGIN", "-o", str(HOST_PROGRAM)])
    return HOST_PROGRAM


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liboracle.so"


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
    print(build_host_program())
    print(build_oracle())
