"""Host-side behaviour of the C ABI (no kernel launches; runs without a GPU).

* libspla.so.1 loads and exports every function declared in include/spla/*.h
  and include/spla_amd/*.h;
* Context defaults, setters and their checks (src/spla/context_internal.hpp:58-140,
  src/spla/context.cpp:122-366) including handle errors;
* MatrixDistribution validation (src/spla/matrix_distribution_internal.cpp:43-150)
  and the Fortran-handle variants;
* parameter errors are raised before any device work;
* MPI: with an initialised one-process MPI the communicator size is queried;
  on two ranks a stripe-stripe-block product is rejected with MPIError.
"""

from __future__ import annotations

import ctypes
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

import spla_amd as spla
from spla_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for hdr in list((ROOT / "include" / "spla").glob("*.h")) + list(
            (ROOT / "include" / "spla_amd").glob("*.h")):
        names.update(re.findall(r"SPLA_EXPORT\s+SplaError\s+(\w+)\s*\(", hdr.read_text()))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = _lib.lib()
    names = declared_functions()
    assert len(names) >= 52
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_cpp_api():
    nm = shutil.which("nm")
    if not nm:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "-C", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    for sym in ("spla::Context::Context(SplaProcessingUnit)",
                "spla::MatrixDistribution::create_blacs_block_cyclic(int, char, int, int, int, int)",
                "spla::gemm(SplaOperation, SplaOperation, int, int, int, std::complex<double>",
                "spla::pgemm_ssb(int, int, int, SplaOperation, double, double const*",
                "spla::pgemm_ssbtr(int, int, int, SplaOperation, std::complex<double>",
                "spla::pgemm_sbs(int, int, int, std::complex<double>"):
        assert sym in out, sym


def test_context_defaults_and_setters():
    ctx = spla.Context(spla.ProcessingUnit.HOST)
    assert ctx.processing_unit() == spla.ProcessingUnit.HOST
    assert ctx.num_tiles() == 4
    assert ctx.tile_size_host() == 500
    assert ctx.tile_size_gpu() == 2048
    assert ctx.op_threshold_gpu() == 2000000
    assert ctx.allocated_memory_host() == 0
    assert ctx.allocated_memory_pinned() == 0
    assert ctx.allocated_memory_gpu() == 0
    ctx.set_num_tiles(7)
    ctx.set_tile_size_host(100)
    ctx.set_tile_size_gpu(512)
    ctx.set_op_threshold_gpu(0)
    assert (ctx.num_tiles(), ctx.tile_size_host(), ctx.tile_size_gpu(),
            ctx.op_threshold_gpu()) == (7, 100, 512, 0)
    for setter, bad in ((ctx.set_num_tiles, 0), (ctx.set_tile_size_host, 0),
                        (ctx.set_tile_size_gpu, -1), (ctx.set_op_threshold_gpu, -1)):
        with pytest.raises(spla.InvalidParameterError):
            setter(bad)


def test_context_handle_errors():
    lib = _lib.lib()
    assert lib.spla_ctx_destroy(None) == spla.SplaError.INVALID_HANDLE_ERROR
    v = ctypes.c_int()
    assert lib.spla_ctx_num_tiles(None, ctypes.byref(v)) == spla.SplaError.INVALID_HANDLE_ERROR
    h = ctypes.c_void_p()
    assert lib.spla_ctx_create(ctypes.byref(h), 7) == spla.SplaError.INVALID_PARAMETER_ERROR
    assert lib.spla_ctx_create(ctypes.byref(h), 0) == 0
    assert lib.spla_ctx_destroy(ctypes.byref(h)) == 0
    assert not h.value  # handle nulled


def test_gpu_context_without_device():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(spla.GPUError):
        spla.Context(spla.ProcessingUnit.GPU)


def test_user_allocators():
    ctx = spla.Context(spla.ProcessingUnit.HOST)
    libc = ctypes.CDLL(None)
    libc.malloc.restype = ctypes.c_void_p
    libc.malloc.argtypes = [ctypes.c_size_t]
    libc.free.argtypes = [ctypes.c_void_p]
    ctx.set_alloc_host(lambda n: libc.malloc(n), lambda p: libc.free(p))
    ctx.set_alloc_pinned(lambda n: libc.malloc(n), lambda p: libc.free(p))
    ctx.set_alloc_gpu(lambda n: libc.malloc(n), lambda p: libc.free(p))
    assert ctx.allocated_memory_host() == 0
    for setter in (ctx.set_alloc_host, ctx.set_alloc_pinned, ctx.set_alloc_gpu):
        with pytest.raises(spla.InvalidAllocatorFunctionError):
            setter(None, lambda p: None)
        with pytest.raises(spla.InvalidAllocatorFunctionError):
            setter(lambda n: 0, None)


def test_block_cyclic_distribution():
    for order in ("R", "r", "C", "c"):
        d = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, order, 1, 1, 64,
                                                              32)
        assert (d.proc_grid_rows(), d.proc_grid_cols(), d.row_block_size(),
                d.col_block_size()) == (1, 1, 64, 32)
        assert d.type() == spla.DistributionType.BLACS_BLOCK_CYCLIC
        assert d.comm() == spla.MPI_COMM_SELF
        d.set_row_block_size(7)
        d.set_col_block_size(9)
        assert (d.row_block_size(), d.col_block_size()) == (7, 9)
    bad = [("X", 1, 1, 64, 64), ("R", 0, 1, 64, 64), ("R", 1, 0, 64, 64), ("R", 1, 1, 0, 64),
           ("R", 1, 1, 64, 0), ("R", 2, 1, 64, 64)]  # 2 x 1 grid needs 2 ranks
    for args in bad:
        with pytest.raises(spla.InvalidParameterError):
            spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, *args)


def test_mapping_and_mirror_distributions():
    d = spla.MatrixDistribution.create_blacs_block_cyclic_from_mapping(spla.MPI_COMM_SELF, [0], 1,
                                                                       1, 16, 16)
    assert d.proc_grid_rows() == 1
    for mapping, rows, cols in (([1], 1, 1), ([0, 0], 2, 1), (None, 1, 1)):
        with pytest.raises(spla.InvalidParameterError):
            spla.MatrixDistribution.create_blacs_block_cyclic_from_mapping(
                spla.MPI_COMM_SELF, mapping, rows, cols, 16, 16)
    m = spla.MatrixDistribution.create_mirror(spla.MPI_COMM_SELF)
    assert m.type() == spla.DistributionType.MIRROR
    assert (m.proc_grid_rows(), m.proc_grid_cols(), m.row_block_size(),
            m.col_block_size()) == (1, 1, 256, 256)


def test_fortran_handle_variants():
    lib = _lib.lib()
    h = ctypes.c_void_p()
    assert lib.spla_mat_dis_create_block_cyclic_fortran(ctypes.byref(h), spla.MPI_COMM_SELF,
                                                        b"C", 1, 1, 8, 8) == 0
    f = ctypes.c_int()
    assert lib.spla_mat_dis_comm_fortran(h, ctypes.byref(f)) == 0
    assert f.value == spla.MPI_COMM_SELF
    assert lib.spla_mat_dis_destroy(ctypes.byref(h)) == 0 and not h.value
    arr = (ctypes.c_int * 1)(0)
    assert lib.spla_mat_dis_create_blacs_block_cyclic_from_mapping_fortran(
        ctypes.byref(h), spla.MPI_COMM_SELF, arr, 1, 1, 8, 8) == 0
    assert lib.spla_mat_dis_destroy(ctypes.byref(h)) == 0
    assert lib.spla_mat_dis_create_mirror_fortran(ctypes.byref(h), spla.MPI_COMM_SELF) == 0
    assert lib.spla_mat_dis_destroy(ctypes.byref(h)) == 0
    assert lib.spla_mat_dis_destroy(None) == spla.SplaError.INVALID_HANDLE_ERROR


def test_parameter_errors_precede_device_work():
    ctx = spla.Context(spla.ProcessingUnit.HOST)
    dist = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_SELF, "R", 1, 1, 64, 64)
    with pytest.raises(spla.InvalidParameterError):
        spla.pgemm_ssb(2, 2, 2, spla.Operation.NONE, 1.0, 1, 2, 1, 2, 0.0, 1, 2, 0, 0, dist, ctx,
                       dtype="complex128")
    with pytest.raises(spla.InvalidParameterError):
        spla.gemm(0, 0, -1, 2, 2, 1.0, 1, 2, 1, 2, 0.0, 1, 2, ctx, dtype="float64")
    with pytest.raises(spla.InvalidPointerError):
        spla.gemm(0, 0, 2, 2, 2, 1.0, None, 2, 1, 2, 0.0, 1, 2, ctx, dtype="float64")
    # m == 0 returns before anything else (no device needed)
    spla.gemm(0, 0, 0, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2, ctx, dtype="float64")


def test_batch_parameter_errors_precede_device_work():
    ctx = spla.Context(spla.ProcessingUnit.HOST)
    good = dict(m=4, n=4, k=4, A=1, lda=4, B=2, ldb=4, C=3, ldc=4)
    # sbs blocks (opA = N): A is m x k (lda >= m) and no fill mode
    with pytest.raises(spla.InvalidParameterError):
        spla.Batch(ctx, [{**good, "m": 5, "ldc": 5}], 1.0, 0.0, opA=spla.Operation.NONE)
    with pytest.raises(spla.InvalidParameterError):
        spla.Batch(ctx, [{**good, "fill": spla.FillMode.LOWER}], 1.0, 0.0,
                   opA=spla.Operation.NONE)
    for bad in (dict(m=-1), dict(lda=3), dict(ldb=2), dict(ldc=3), dict(fill=7),
                dict(row_offset=-1)):
        with pytest.raises(spla.InvalidParameterError):
            spla.Batch(ctx, [{**good, **bad}], 1.0, 0.0)
    with pytest.raises(spla.InvalidPointerError):
        spla.Batch(ctx, [{**good, "A": None}], 1.0, 0.0)
    lib = _lib.lib()
    h = ctypes.c_void_p()
    one = (ctypes.c_double * 2)(1.0, 0.0)
    assert lib.spla_amd_batch_create_ssb(ctypes.byref(h), ctx.handle, 5, 2, 0, None, one, one,
                                         0) == spla.SplaError.INVALID_PARAMETER_ERROR
    assert lib.spla_amd_batch_create_ssb(ctypes.byref(h), None, 1, 2, 0, None, one, one,
                                         0) == spla.SplaError.INVALID_HANDLE_ERROR
    assert lib.spla_amd_batch_create_ssb(ctypes.byref(h), ctx.handle, 1, 2, 1, None, one, one,
                                         0) == spla.SplaError.INVALID_POINTER_ERROR
    assert lib.spla_amd_batch_execute(None) == spla.SplaError.INVALID_HANDLE_ERROR
    v = ctypes.c_int()
    assert lib.spla_amd_batch_split_blocks(None, ctypes.byref(v)) == \
        spla.SplaError.INVALID_HANDLE_ERROR
    assert lib.spla_amd_batch_destroy(None) == spla.SplaError.INVALID_HANDLE_ERROR


MPI_LIB = "/opt/conda/lib/libmpi.so"
MPIEXEC = "/opt/conda/bin/mpiexec"

MPI_SCRIPT = r"""
import ctypes, sys
sys.path.insert(0, {root!r})
mpi = ctypes.CDLL({lib!r}, mode=ctypes.RTLD_GLOBAL)
mpi.MPI_Init(None, None)
import spla_amd as spla
size = ctypes.c_int()
mpi.MPI_Comm_size(spla.MPI_COMM_WORLD, ctypes.byref(size))
ctx = spla.Context(spla.ProcessingUnit.HOST)
d = spla.MatrixDistribution.create_blacs_block_cyclic(spla.MPI_COMM_WORLD, "R", size.value, 1,
                                                      16, 16)
assert d.proc_grid_rows() == size.value
status = "local"
try:
    spla.pgemm_ssb(2, 2, 2, 2, 1.0, 1, 2, 1, 2, 0.0, 1, 2, 0, 0, d, ctx, dtype="complex128")
except spla.MPIError:
    status = "mpierror"
except spla.GPUError:
    status = "gpu"
print("RESULT", size.value, status, flush=True)
mpi.MPI_Finalize()
"""


@pytest.mark.skipif(not (os.path.exists(MPI_LIB) and os.path.exists(MPIEXEC)),
                    reason="MPICH not available")
@pytest.mark.parametrize("ranks", [1, 2])
def test_mpi_communicator_size(tmp_path, ranks):
    script = tmp_path / "mpi_case.py"
    script.write_text(MPI_SCRIPT.format(root=str(ROOT), lib=MPI_LIB))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="", ROCR_VISIBLE_DEVICES="")
    res = subprocess.run([MPIEXEC, "-n", str(ranks), sys.executable, str(script)],
                         capture_output=True, text=True, timeout=120, env=env)
    lines = [l for l in res.stdout.splitlines() if l.startswith("RESULT")]
    assert len(lines) == ranks, res.stdout + res.stderr
    for line in lines:
        _, size, status = line.split()
        assert int(size) == ranks
        # one rank: the product reaches the device engine (no GPU here -> GPU error);
        # two ranks: the single-rank engine rejects the communicator
        assert status == ("mpierror" if ranks > 1 else "gpu"), line
