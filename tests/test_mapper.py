"""The C++ mapper layer (include/rvk_mapper.h) through its compiled checks.

tests/cpp/rvk_mapper_test is built by robartsvtk_amd/build.py next to
librvkhip.so.  Its cpu mode covers the transfer-function objects (piecewise
functions, polygon rasterisation against closed forms, the .2tf format) and the
mapper error paths; its gpu mode renders through the DRR, 1D and 2D mappers and
compares with direct C-ABI renders of the same view (DRR bit-exact, RGBA8
within 1 LSB on >= 99.9% of channels), checks the voxel-space clipping planes
against the world-space ones, multi-frame input, a device move with data and a
far z-buffer.
"""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "rvk_mapper_test"


def run(mode: str, timeout: int):
    if not BIN.exists():
        pytest.skip(f"{BIN} is not built (python robartsvtk_amd/build.py)")
    res = subprocess.run([str(BIN), mode], capture_output=True, text=True, timeout=timeout)
    print(res.stdout)
    print(res.stderr)
    fails = [ln for ln in res.stdout.splitlines() if ln.startswith("FAIL")]
    passes = [ln for ln in res.stdout.splitlines() if ln.startswith("PASS")]
    assert res.returncode == 0 and not fails, fails or res.stderr
    return passes


def test_mapper_host_logic():
    passes = run("cpu", 60)
    assert len(passes) >= 10


@pytest.mark.gpu
def test_mapper_renders_match_direct_calls():
    passes = run("gpu", 110)
    assert len(passes) >= 10
