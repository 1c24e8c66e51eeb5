"""Build recipe of libspla.so.1 (HIP, gfx950) without CMake.

Every translation unit under ``spla_amd/csrc`` is compiled with hipcc for
gfx950 (in parallel) into ``spla_amd/_build`` and linked into
``spla_amd/lib/libspla.so.1`` (SONAME ``libspla.so.1`` as SPLA's SOVERSION 1)
plus a ``libspla.so`` link for ``-lspla``. Run ``python -m spla_amd.build``.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "lib"
LIBNAME = "libspla.so.1"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
MPI_INCLUDE = "/opt/conda/include"

COMMON = [
    "-std=c++17",
    "-O3",
    "-fPIC",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-function",
    f"--offload-arch={ARCH}",
    "-I", str(CSRC),
    "-I", str(ROOT / "include"),
    "-idirafter", MPI_INCLUDE,
    # packed FP32 instructions are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources() -> list[Path]:
    return sorted(p for p in CSRC.rglob("*") if p.suffix in (".cpp", ".hip"))


def _compile(src: Path) -> tuple[Path, str]:
    obj = BUILD / (str(src.relative_to(CSRC)).replace("/", "__") + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        line for line in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"compilation of {src} failed:\n{out}")
    return obj, out


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compiles every source (always) and links the library. Returns its path."""
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 1, len(srcs))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, out in results:
            if out.strip():
                print(out)
    target = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [
        HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC",
        "-o", str(tmp), *map(str, objs),
        f"-Wl,-soname,{LIBNAME}", "-Wl,--no-undefined", "-ldl", "-lpthread",
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}{res.stderr}")
    os.replace(tmp, target)
    link = LIBDIR / "libspla.so"
    if link.is_symlink() or link.exists():
        link.unlink()
    link.symlink_to(LIBNAME)
    return target


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
