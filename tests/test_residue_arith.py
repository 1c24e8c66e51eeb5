"""The residue arithmetic of the split-precision kernels, run on the CPU.

tests/cpp/residue_arith_check.cpp compiles spla_amd/csrc/kernels/residue_arith.hpp
(the same header the residue kernels include) with g++ and compares the
symmetric residue of every value with exact integer arithmetic: values around
every power of two up to 2^53 and 2e6 random integers, all 20 moduli.
"""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_residue_arithmetic_is_exact(tmp_path):
    exe = tmp_path / "residue_arith_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off",
                    "-I", str(ROOT / "spla_amd" / "csrc"),
                    str(ROOT / "tests" / "cpp" / "residue_arith_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 mismatches" in out.stdout
