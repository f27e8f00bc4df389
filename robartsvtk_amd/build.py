"""Build the gfx950 library librvkhip.so in-tree with hipcc.

Every source under csrc/ is compiled to an object in parallel (HIP sources
for --offload-arch=gfx950, host C++ sources with the same compiler), then
linked into robartsvtk_amd/librvkhip.so next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "librvkhip.so"

ARCH = os.environ.get("RVK_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-Wno-unused-value", "-I", str(PKG.parent / "include")]
# Packed FP32 (v_pk_*_f32) is kept out of the device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")) + list(CSRC.glob("mapper/*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.relative_to(CSRC).as_posix().replace("/", "__") + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-x", "c++", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build_library(verbose: bool = False, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


TESTS = PKG.parent / "tests" / "cpp"
MAPPER_TEST = TESTS / "rvk_mapper_test"


def build_tests(verbose: bool = False) -> Path:
    """The C++ mapper checks (tests/cpp/mapper_test.cpp), linked against the
    library with an rpath relative to the executable."""
    cmd = [HIPCC, *COMMON, "-x", "c++", str(TESTS / "mapper_test.cpp"), "-x", "none", "-o", str(MAPPER_TEST),
           f"-L{PKG}", "-lrvkhip", "-Wl,-rpath,$ORIGIN/../../robartsvtk_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"test build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {MAPPER_TEST}", file=sys.stderr)
    return MAPPER_TEST


if __name__ == "__main__":
    build_library(verbose=True)
    build_tests(verbose=True)
