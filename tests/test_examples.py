"""C and C++ programs written against include/spla (tests/examples) link
against libspla.so.1 and run: a C program calling spla_pzgemm_ssb and a C++
program calling spla::pgemm_sbs / spla::pgemm_ssbtr (SPLA's examples/example.c
and examples/example.cpp use the same API). They are built by
__graft_entry__.build() (tests/examples/Makefile)."""

import subprocess
from pathlib import Path

import pytest

BUILD = Path(__file__).resolve().parent / "examples" / "_build"


def _exe(name):
    exe = BUILD / name
    if not exe.exists():
        pytest.fail(f"{exe} not built: run `make -C tests/examples` (part of build())")
    return exe


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except ImportError:
        return False


@pytest.mark.parametrize("name", ["example_ssb", "example_sbs"])
def test_examples_link_and_report_missing_device(name):
    if _has_gpu():
        pytest.skip("a GPU is present (the GPU test runs the programs)")
    res = subprocess.run([str(_exe(name))], capture_output=True, text=True, timeout=120)
    # without a GPU the context creation fails cleanly: SplaError 13 / GPUNoDeviceError
    assert res.returncode != 0
    assert "13" in res.stdout + res.stderr or "GPU" in res.stdout + res.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name,args", [("example_ssb", []), ("example_ssb", ["300", "17", "40000"]),
                                       ("example_sbs", [])])
def test_examples_run_on_gpu(name, args):
    res = subprocess.run([str(_exe(name)), *args], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "max error" in res.stdout
