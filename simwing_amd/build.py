"""Builds the in-tree native libraries.

* ``simwing_amd/libsimwing_xpbd.so`` -- the product: host C++ (definition,
  assembly, schedules, C-ABI) and the gfx950 HIP kernels, compiled with
  hipcc for ``--offload-arch=gfx950`` with floating-point contraction off.
* ``oracle/liboracle_xpbd.so`` -- the CPU oracle (test infrastructure),
  compiled with g++ by ``oracle/Makefile``.

Run ``python -m simwing_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
BUILD = PACKAGE / "_build"
LIBRARY = PACKAGE / "libsimwing_xpbd.so"
ORACLE_LIBRARY = ROOT / "oracle" / "liboracle_xpbd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SIMWING_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-std=c++20",
    "-O3",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wno-unused-function",
    "-I", str(ROOT / "include"),
]
DEVICE = [
    "-x", "hip",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]

SOURCES = [
    ("xpbd_device.hip", True),
    ("capi.cpp", False),
    ("definition.cpp", False),
    ("assembly.cpp", False),
    ("schedule.cpp", False),
    ("frame.cpp", False),
    ("scene_io.cpp", False),
    ("wing_generator.cpp", False),
    ("checkpoint_codec.cpp", False),
]


def _compile(source: str, device: bool) -> Path:
    out = BUILD / (Path(source).stem + ".o")
    cmd = [HIPCC, *COMMON, *(DEVICE if device else []), "-c", str(CSRC / source), "-o", str(out)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{result.stdout}\n{result.stderr}")
    return out


def build_product(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        objects = list(pool.map(lambda s: _compile(*s), SOURCES))
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(LIBRARY), *map(str, objects)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{result.stdout}\n{result.stderr}")
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


def build_oracle(verbose: bool = False, reference: bool | None = None) -> Path:
    result = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-j8"], capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{result.stdout}\n{result.stderr}")
    if reference is None:
        reference = Path("/root/reference/src/softwing").is_dir()
    if reference:
        result = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-j8", "ref"], capture_output=True, text=True)
        if result.returncode != 0:
            raise RuntimeError(f"reference build failed\n{result.stdout}\n{result.stderr}")
    if verbose:
        print(f"built {ORACLE_LIBRARY}")
    return ORACLE_LIBRARY


def build(verbose: bool = False) -> None:
    build_product(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
