"""Host test of the device pool allocator's caching policy (no GPU needed).

Builds tests/cpp/test_pool_allocator.cpp with g++ against the header-only pool
(spla_amd/csrc/runtime/pool_allocator.hpp) and runs it; the program mirrors
SPLA's tests/test_pool_allocator.cpp:17-109 (no leak, no double free, every
block at least the requested size) and adds the retry-on-failure path.
"""

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_pool_allocator(tmp_path):
    exe = tmp_path / "test_pool_allocator"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fsanitize=address,undefined",
                    "-I", str(ROOT / "spla_amd" / "csrc"),
                    str(ROOT / "tests" / "cpp" / "test_pool_allocator.cpp"), "-o", str(exe)],
                   check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "OK" in res.stdout
