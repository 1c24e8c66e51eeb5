"""The integer arithmetic of the split-precision kernels, run on the CPU.

* tests/cpp/residue_arith_check.cpp compiles spla_amd/csrc/kernels/residue_arith.hpp
  (the same header the residue kernels include) with g++ and compares the
  residues of every value (all three planes of a complex entry, as the kernels
  form them) with exact integer arithmetic: values around every power of two
  up to 2^53 and 2e6 random pairs, all 20 moduli;
* tests/cpp/crt_arith_check.cpp compiles spla_amd/csrc/kernels/crt_arith.hpp (the
  reconstruction of the CRT kernel) with the planner's constants and compares
  every reconstruction with exact 192-bit integers, for 1..20 moduli, the edges
  of the canonical range, the complex combination from biased bytes and from
  INT32 split sums, and the symmetric residue the GEMM epilogue stores.
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


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_crt_reconstruction_is_exact(tmp_path):
    exe = tmp_path / "crt_arith_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off",
                    "-I", str(ROOT / "spla_amd" / "csrc"),
                    str(ROOT / "spla_amd" / "csrc" / "runtime" / "split_plan.cpp"),
                    str(ROOT / "tests" / "cpp" / "crt_arith_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " 0 mismatches" in out.stdout
