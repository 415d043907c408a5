"""Build recipe for the gfx950 library (and the oracle used by the tests).

``python -m imops_amd.build`` compiles every HIP/C++ source under
``imops_amd/csrc`` for gfx950 with hipcc into ``imops_amd/libimops_gpu.so``
(in-tree, so it travels with the repository snapshot) and builds
``oracle/liboracle.so`` with make.

Device code is compiled with -ffp-contract=off (no FMA contraction: the
reference's Rust never fuses a*b+c), correctly rounded f32 division/sqrt, and
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libimops_gpu.so"
ORACLE = ROOT / "oracle"

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", f"{ROCM}/bin/hipcc")
ARCH = os.environ.get("IMOPS_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
HIP_FLAGS = COMMON + [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# host-only C++ without HIP calls (OpenMP parallel generator)
GXX_FLAGS = COMMON + ["-fopenmp", "-fno-fast-math"]
GXX_ONLY = {"synth.cpp"}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.name in GXX_ONLY:
        cmd = ["g++", *GXX_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_library(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs), "-lgomp",
          f"-L{ROCM}/lib", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    _run(["make", "-C", str(ORACLE), "-j8", "-s"])
    return ORACLE / "liboracle.so"


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


def main(argv: list[str]) -> int:
    if "--clean" in argv:
        clean()
    lib = build_library()
    print(f"built {lib}")
    if "--no-oracle" not in argv:
        print(f"built {build_oracle()}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
