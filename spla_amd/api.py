"""Python mirror of SPLA's API over the C ABI of libspla.so.1.

Names, argument order and meaning follow SPLA's C++ API
(include/spla/{context,matrix_distribution,gemm,pgemm_ssb,pgemm_ssbtr,pgemm_sbs}.hpp);
errors are raised as the exception class of the returned SplaError code
(include/spla/exceptions.hpp:41-172). Arrays are numpy arrays (host memory)
or torch tensors (host or device memory), column-major with explicit leading
dimensions, exactly as the C pointers they stand for; plain integers are taken
as raw addresses. The element type selects the s/d/c/z entry point.
"""

from __future__ import annotations

import ctypes
import enum

import numpy as np

from spla_amd import _lib


class ProcessingUnit(enum.IntEnum):
    HOST = 0
    GPU = 1


class Operation(enum.IntEnum):
    NONE = 0
    TRANSPOSE = 1
    CONJ_TRANSPOSE = 2


class FillMode(enum.IntEnum):
    FULL = 0
    UPPER = 1
    LOWER = 2


class GemmMode(enum.IntEnum):
    """Arithmetic of double / complex-double products (include/spla_amd/spla_amd.h)."""
    FP64 = 0
    SPLIT_INT8 = 1
    AUTO = 2


class DistributionType(enum.IntEnum):
    BLACS_BLOCK_CYCLIC = 0
    MIRROR = 1


class SplaError(enum.IntEnum):
    SUCCESS = 0
    UNKNOWN_ERROR = 1
    INTERNAL_ERROR = 2
    INVALID_PARAMETER_ERROR = 3
    INVALID_POINTER_ERROR = 4
    INVALID_HANDLE_ERROR = 5
    MPI_ERROR = 6
    MPI_ALLOCATION_ERROR = 7
    MPI_THREAD_SUPPORT_ERROR = 8
    GPU_ERROR = 9
    GPU_SUPPORT_ERROR = 10
    GPU_ALLOCATION_ERROR = 11
    GPU_LAUNCH_ERROR = 12
    GPU_NO_DEVICE_ERROR = 13
    GPU_INVALID_VALUE_ERROR = 14
    GPU_INVALID_DEVICE_POINTER_ERROR = 15
    GPU_BLAS_ERROR = 16
    INVALID_ALLOCATOR_FUNCTION = 17


class GenericError(Exception):
    code = SplaError.UNKNOWN_ERROR


def _err(name, base, code):
    return type(name, (base,), {"code": code})


InternalError = _err("InternalError", GenericError, SplaError.INTERNAL_ERROR)
InvalidParameterError = _err("InvalidParameterError", GenericError,
                             SplaError.INVALID_PARAMETER_ERROR)
InvalidPointerError = _err("InvalidPointerError", GenericError, SplaError.INVALID_POINTER_ERROR)
InvalidHandleError = _err("InvalidHandleError", GenericError, SplaError.INVALID_HANDLE_ERROR)
InvalidAllocatorFunctionError = _err("InvalidAllocatorFunctionError", GenericError,
                                     SplaError.INVALID_ALLOCATOR_FUNCTION)
MPIError = _err("MPIError", GenericError, SplaError.MPI_ERROR)
MPIAllocError = _err("MPIAllocError", MPIError, SplaError.MPI_ALLOCATION_ERROR)
MPIThreadSupportError = _err("MPIThreadSupportError", MPIError,
                             SplaError.MPI_THREAD_SUPPORT_ERROR)
GPUError = _err("GPUError", GenericError, SplaError.GPU_ERROR)
GPUSupportError = _err("GPUSupportError", GPUError, SplaError.GPU_SUPPORT_ERROR)
GPUAllocationError = _err("GPUAllocationError", GPUError, SplaError.GPU_ALLOCATION_ERROR)
GPULaunchError = _err("GPULaunchError", GPUError, SplaError.GPU_LAUNCH_ERROR)
GPUNoDeviceError = _err("GPUNoDeviceError", GPUError, SplaError.GPU_NO_DEVICE_ERROR)
GPUInvalidValueError = _err("GPUInvalidValueError", GPUError, SplaError.GPU_INVALID_VALUE_ERROR)
GPUInvalidDevicePointerError = _err("GPUInvalidDevicePointerError", GPUError,
                                    SplaError.GPU_INVALID_DEVICE_POINTER_ERROR)
GPUBlasError = _err("GPUBlasError", GPUError, SplaError.GPU_BLAS_ERROR)

_BY_CODE = {cls.code: cls for cls in (
    InternalError, InvalidParameterError, InvalidPointerError, InvalidHandleError,
    InvalidAllocatorFunctionError, MPIError, MPIAllocError, MPIThreadSupportError, GPUError,
    GPUSupportError, GPUAllocationError, GPULaunchError, GPUNoDeviceError, GPUInvalidValueError,
    GPUInvalidDevicePointerError, GPUBlasError)}


def check(code: int) -> None:
    """Raises the exception class of a non-zero SplaError code."""
    if code != 0:
        cls = _BY_CODE.get(code, GenericError)
        raise cls(f"SPLA error {SplaError(code).name if code in SplaError._value2member_map_ else code}")


MPI_COMM_WORLD = _lib.MPI_COMM_WORLD
MPI_COMM_SELF = _lib.MPI_COMM_SELF


class Context:
    """spla::Context (include/spla/context.hpp:51-278)."""

    def __init__(self, pu: ProcessingUnit = ProcessingUnit.GPU):
        self._lib = _lib.lib()
        self._h = ctypes.c_void_p()
        check(self._lib.spla_ctx_create(ctypes.byref(self._h), int(pu)))
        self._keep = []  # user allocator callbacks stay alive with the context

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            check(self._lib.spla_ctx_destroy(ctypes.byref(self._h)))

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass

    def _get(self, name, ctype=ctypes.c_int):
        v = ctype()
        check(getattr(self._lib, f"spla_ctx_{name}")(self._h, ctypes.byref(v)))
        return v.value

    def processing_unit(self) -> ProcessingUnit:
        return ProcessingUnit(self._get("processing_unit"))

    def num_tiles(self) -> int:
        return self._get("num_tiles")

    def tile_size_host(self) -> int:
        return self._get("tile_size_host")

    def tile_size_gpu(self) -> int:
        return self._get("tile_size_gpu")

    def op_threshold_gpu(self) -> int:
        return self._get("op_threshold_gpu")

    def gpu_device_id(self) -> int:
        return self._get("gpu_device_id")

    def allocated_memory_host(self) -> int:
        return self._get("allocated_memory_host", ctypes.c_uint64)

    def allocated_memory_pinned(self) -> int:
        return self._get("allocated_memory_pinned", ctypes.c_uint64)

    def allocated_memory_gpu(self) -> int:
        return self._get("allocated_memory_gpu", ctypes.c_uint64)

    def set_num_tiles(self, v: int) -> None:
        check(self._lib.spla_ctx_set_num_tiles(self._h, v))

    def set_tile_size_host(self, v: int) -> None:
        check(self._lib.spla_ctx_set_tile_size_host(self._h, v))

    def set_tile_size_gpu(self, v: int) -> None:
        check(self._lib.spla_ctx_set_tile_size_gpu(self._h, v))

    def set_op_threshold_gpu(self, v: int) -> None:
        check(self._lib.spla_ctx_set_op_threshold_gpu(self._h, v))

    def _set_alloc(self, which, allocate, deallocate):
        fn = getattr(self._lib, f"spla_ctx_set_alloc_{which}")
        alloc_t, free_t = fn.argtypes[1], fn.argtypes[2]
        a = alloc_t(allocate) if allocate is not None else ctypes.cast(None, alloc_t)
        f = free_t(deallocate) if deallocate is not None else ctypes.cast(None, free_t)
        self._keep.append((a, f))
        check(fn(self._h, a, f))

    def set_alloc_host(self, allocate, deallocate) -> None:
        self._set_alloc("host", allocate, deallocate)

    def set_alloc_pinned(self, allocate, deallocate) -> None:
        self._set_alloc("pinned", allocate, deallocate)

    def set_alloc_gpu(self, allocate, deallocate) -> None:
        self._set_alloc("gpu", allocate, deallocate)

    # -- extensions (include/spla_amd/spla_amd.h) --
    def set_stream(self, stream) -> None:
        """Runs the kernels on a caller stream (torch.cuda.Stream, raw handle or None)."""
        handle = getattr(stream, "cuda_stream", stream)
        check(self._lib.spla_amd_ctx_set_stream(self._h, handle or None))

    def set_gemm_mode(self, mode: GemmMode) -> None:
        check(self._lib.spla_amd_ctx_set_gemm_mode(self._h, int(mode)))

    def set_split_tolerance(self, tol: float) -> None:
        check(self._lib.spla_amd_ctx_set_split_tolerance(self._h, float(tol)))

    def set_split_workspace_limit(self, nbytes: int) -> None:
        """Largest split-precision workspace (bytes; 0: a quarter of the device memory)."""
        check(self._lib.spla_amd_ctx_set_split_workspace_limit(self._h, int(nbytes)))

    def last_product_info(self) -> dict:
        """Form of the last double / complex-double product (split plan, error bound)."""
        v = [ctypes.c_int() for _ in range(5)]
        b = ctypes.c_double()
        check(self._lib.spla_amd_ctx_last_product_info(self._h, *[ctypes.byref(x) for x in v],
                                                       ctypes.byref(b)))
        return {"split": bool(v[0].value), "reason": v[1].value, "moduli": v[2].value,
                "bits_a": v[3].value, "bits_b": v[4].value, "bound": b.value}

    PHASES = ("gemm_fp64", "splitk_reduce", "split_stats", "split_residues", "split_int8_gemm",
              "split_crt")

    def set_timing(self, enable: bool) -> None:
        check(self._lib.spla_amd_ctx_set_timing(self._h, 1 if enable else 0))

    def last_timing(self) -> dict:
        """Milliseconds per kernel phase of the last call (timing must be enabled)."""
        ms = (ctypes.c_double * len(self.PHASES))()
        check(self._lib.spla_amd_ctx_last_timing(self._h, ms))
        return dict(zip(self.PHASES, list(ms)))

    def set_stream_chunk_bytes(self, nbytes: int) -> None:
        """Target bytes of one streamed k-chunk of host-resident operands."""
        check(self._lib.spla_amd_ctx_set_stream_chunk_bytes(self._h, int(nbytes)))


class MatrixDistribution:
    """spla::MatrixDistribution (include/spla/matrix_distribution.hpp:47-222)."""

    def __init__(self, handle):
        self._lib = _lib.lib()
        self._h = handle

    @classmethod
    def create_blacs_block_cyclic(cls, comm: int, order: str, proc_grid_rows: int,
                                  proc_grid_cols: int, row_block_size: int,
                                  col_block_size: int) -> "MatrixDistribution":
        h = ctypes.c_void_p()
        check(_lib.lib().spla_mat_dis_create_block_cyclic(
            ctypes.byref(h), comm, order.encode()[:1], proc_grid_rows, proc_grid_cols,
            row_block_size, col_block_size))
        return cls(h)

    @classmethod
    def create_blacs_block_cyclic_from_mapping(cls, comm: int, mapping, proc_grid_rows: int,
                                               proc_grid_cols: int, row_block_size: int,
                                               col_block_size: int) -> "MatrixDistribution":
        h = ctypes.c_void_p()
        arr = None
        if mapping is not None:
            arr = (ctypes.c_int * len(mapping))(*mapping)
        check(_lib.lib().spla_mat_dis_create_blacs_block_cyclic_from_mapping(
            ctypes.byref(h), comm, arr, proc_grid_rows, proc_grid_cols, row_block_size,
            col_block_size))
        return cls(h)

    @classmethod
    def create_mirror(cls, comm: int) -> "MatrixDistribution":
        h = ctypes.c_void_p()
        check(_lib.lib().spla_mat_dis_create_mirror(ctypes.byref(h), comm))
        return cls(h)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            check(self._lib.spla_mat_dis_destroy(ctypes.byref(self._h)))

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass

    def _get(self, name, ctype=ctypes.c_int):
        v = ctype()
        check(getattr(self._lib, f"spla_mat_dis_{name}")(self._h, ctypes.byref(v)))
        return v.value

    def proc_grid_rows(self) -> int:
        return self._get("proc_grid_rows")

    def proc_grid_cols(self) -> int:
        return self._get("proc_grid_cols")

    def row_block_size(self) -> int:
        return self._get("row_block_size")

    def col_block_size(self) -> int:
        return self._get("col_block_size")

    def type(self) -> DistributionType:
        return DistributionType(self._get("type"))

    def comm(self) -> int:
        return self._get("comm")

    def set_row_block_size(self, v: int) -> None:
        check(self._lib.spla_mat_dis_set_row_block_size(self._h, v))

    def set_col_block_size(self, v: int) -> None:
        check(self._lib.spla_mat_dis_set_col_block_size(self._h, v))


class Batch:
    """Batched ring-step block products (spla_amd_batch_* of include/spla_amd/spla_amd.h).

    blocks: sequence of dicts with keys m, n, k, A, lda, B, ldb, C, ldc and optional
    fill (FillMode), row_offset, col_offset; arrays are device tensors or raw
    device addresses. opA TRANSPOSE / CONJ_TRANSPOSE: pgemm_ssb blocks (A k x m);
    opA NONE: pgemm_sbs blocks (A m x k, spla_amd_batch_create_sbs).
    """

    def __init__(self, ctx: Context, blocks, alpha, beta, complex_: bool = True,
                 opA: Operation = Operation.CONJ_TRANSPOSE, measure: bool = False,
                 persistent: bool = False):
        self._lib = ctx._lib
        self._ctx = ctx  # keeps the context alive
        arr = (_lib.SplaAmdBlock * max(1, len(blocks)))()
        for i, b in enumerate(blocks):
            arr[i] = _lib.SplaAmdBlock(b["m"], b["n"], b["k"], _ptr(b["A"]), b["lda"],
                                       _ptr(b["B"]), b["ldb"], _ptr(b["C"]), b["ldc"],
                                       int(b.get("fill", FillMode.FULL)),
                                       b.get("row_offset", 0), b.get("col_offset", 0))
        dt = np.complex128 if complex_ else np.float64
        self._alpha = np.array([alpha], dtype=dt)
        self._beta = np.array([beta], dtype=dt)
        self._h = ctypes.c_void_p()
        tuning = 1 if measure else (2 if persistent else 0)  # SplaAmdTuning
        if int(opA) == int(Operation.NONE):
            check(self._lib.spla_amd_batch_create_sbs(
                ctypes.byref(self._h), ctx.handle, 1 if complex_ else 0, len(blocks), arr,
                self._alpha.ctypes.data, self._beta.ctypes.data, tuning))
        else:
            check(self._lib.spla_amd_batch_create_ssb(
                ctypes.byref(self._h), ctx.handle, 1 if complex_ else 0, int(opA), len(blocks),
                arr, self._alpha.ctypes.data, self._beta.ctypes.data, tuning))

    def execute(self) -> None:
        check(self._lib.spla_amd_batch_execute(self._h))

    def info(self) -> dict:
        cfg, items, wgs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        model, meas = ctypes.c_double(), ctypes.c_double()
        check(self._lib.spla_amd_batch_info(self._h, ctypes.byref(cfg), ctypes.byref(items),
                                            ctypes.byref(wgs), ctypes.byref(model),
                                            ctypes.byref(meas)))
        pers = ctypes.c_int()
        check(self._lib.spla_amd_batch_persistent(self._h, ctypes.byref(pers)))
        split = ctypes.c_int()
        check(self._lib.spla_amd_batch_split_blocks(self._h, ctypes.byref(split)))
        return {"tile_config": cfg.value, "items": items.value, "workgroups": wgs.value,
                "model_ms": model.value, "measured_ms": meas.value,
                "persistent": bool(pers.value), "split_blocks": split.value}

    def close(self) -> None:
        if self._h:
            check(self._lib.spla_amd_batch_destroy(ctypes.byref(self._h)))

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass


def ssb_ring_blocks(dist_type, grid_rows, grid_cols, comm_size, m, n, row_offset, col_offset,
                    fill, tile_size_host):
    """spla_amd_ssb_ring_blocks: (rows_in_block, cols_in_block, [(row, col, rows, cols)])."""
    lib = _lib.lib()
    rib, cib, cnt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    args = (int(dist_type), grid_rows, grid_cols, comm_size, m, n, row_offset, col_offset,
            int(fill), tile_size_host, ctypes.byref(rib), ctypes.byref(cib), ctypes.byref(cnt))
    check(lib.spla_amd_ssb_ring_blocks(*args, None, 0))
    arr = (_lib.SplaAmdRingBlock * max(1, cnt.value))()
    check(lib.spla_amd_ssb_ring_blocks(*args, arr, cnt.value))
    return rib.value, cib.value, [(b.row, b.col, b.numRows, b.numCols) for b in arr[:cnt.value]]


def region_pieces(dist_type, rows_in_block, cols_in_block, grid_rows, grid_cols, rows, cols,
                  row_offset, col_offset):
    """spla_amd_region_pieces: list of 9-tuples (globalRow, globalCol, subRow, subCol,
    localRow, localCol, numRows, numCols, rank)."""
    lib = _lib.lib()
    cnt = ctypes.c_int()
    args = (int(dist_type), rows_in_block, cols_in_block, grid_rows, grid_cols, rows, cols,
            row_offset, col_offset, ctypes.byref(cnt))
    check(lib.spla_amd_region_pieces(*args, None, 0))
    arr = (_lib.SplaAmdPiece * max(1, cnt.value))()
    check(lib.spla_amd_region_pieces(*args, arr, cnt.value))
    return [tuple(getattr(q, f) for f, _ in _lib.SplaAmdPiece._fields_) for q in arr[:cnt.value]]


# ---------------------------------------------------------------------------
# array plumbing

_PREFIX = {np.dtype(np.float32): "s", np.dtype(np.float64): "d",
           np.dtype(np.complex64): "c", np.dtype(np.complex128): "z"}


class timer:  # noqa: N801 - mirrors SPLA's spla_timer_* family
    """Process-wide wall-clock timer (SPLA src/timing/timing.cpp:49-86)."""

    @staticmethod
    def start(name: str) -> None:
        b = name.encode()
        check(_lib.lib().spla_timer_start(len(b), b))

    @staticmethod
    def stop(name: str) -> None:
        b = name.encode()
        check(_lib.lib().spla_timer_stop(len(b), b))

    @staticmethod
    def export_json(path: str) -> None:
        b = str(path).encode()
        check(_lib.lib().spla_timer_export_json(len(b), b))

    @staticmethod
    def print() -> None:
        check(_lib.lib().spla_timer_print())

    @staticmethod
    def set_library_scopes(enable: bool) -> None:
        check(_lib.lib().spla_amd_timer_set_library_scopes(1 if enable else 0))

    @staticmethod
    def clear() -> None:
        check(_lib.lib().spla_amd_timer_clear())


def _dtype_of(x):
    if isinstance(x, np.ndarray):
        return x.dtype
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return {torch.float32: np.dtype(np.float32), torch.float64: np.dtype(np.float64),
                    torch.complex64: np.dtype(np.complex64),
                    torch.complex128: np.dtype(np.complex128)}[x.dtype]
    except ImportError:  # pragma: no cover
        pass
    return None


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data if x.size else None
    data_ptr = getattr(x, "data_ptr", None)
    if data_ptr is not None:
        return data_ptr() or None
    raise TypeError(f"unsupported array type {type(x)}")


def _prefix(dtype, *arrays):
    if dtype is None:
        for a in arrays:
            dt = _dtype_of(a)
            if dt is not None:
                dtype = dt
                break
    dtype = np.dtype(dtype if dtype is not None else np.float64)
    if dtype not in _PREFIX:
        raise TypeError(f"unsupported element type {dtype}")
    return _PREFIX[dtype], dtype


def _scalar(prefix, v):
    if prefix == "s":
        return ctypes.c_float(float(np.real(v)))
    if prefix == "d":
        return ctypes.c_double(float(np.real(v)))
    ct = np.complex64 if prefix == "c" else np.complex128
    buf = np.array([v], dtype=ct)
    return buf  # kept alive by the caller frame


def _scalar_arg(prefix, buf):
    return buf if prefix in "sd" else ctypes.c_void_p(buf.ctypes.data)


def gemm(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ctx: Context, dtype=None):
    """spla::gemm (include/spla/gemm.hpp:66-100): C <- alpha op(A) op(B) + beta C."""
    p, _ = _prefix(dtype, C, A, B)
    a, b = _scalar(p, alpha), _scalar(p, beta)
    check(getattr(ctx._lib, f"spla_{p}gemm")(
        int(opA), int(opB), m, n, k, _scalar_arg(p, a), _ptr(A), lda, _ptr(B), ldb,
        _scalar_arg(p, b), _ptr(C), ldc, ctx.handle))


def pgemm_ssb(m, n, kLocal, opA, alpha, A, lda, B, ldb, beta, C, ldc, cRowOffset, cColOffset,
              distC: MatrixDistribution, ctx: Context, dtype=None):
    """spla::pgemm_ssb (include/spla/pgemm_ssb.hpp:88-120)."""
    p, _ = _prefix(dtype, C, A, B)
    a, b = _scalar(p, alpha), _scalar(p, beta)
    check(getattr(ctx._lib, f"spla_p{p}gemm_ssb")(
        m, n, kLocal, int(opA), _scalar_arg(p, a), _ptr(A), lda, _ptr(B), ldb, _scalar_arg(p, b),
        _ptr(C), ldc, cRowOffset, cColOffset, distC.handle, ctx.handle))


def pgemm_ssbtr(m, n, kLocal, opA, alpha, A, lda, B, ldb, beta, C, ldc, cRowOffset, cColOffset,
                cFillMode, distC: MatrixDistribution, ctx: Context, dtype=None):
    """spla::pgemm_ssbtr (include/spla/pgemm_ssbtr.hpp:94-128)."""
    p, _ = _prefix(dtype, C, A, B)
    a, b = _scalar(p, alpha), _scalar(p, beta)
    check(getattr(ctx._lib, f"spla_p{p}gemm_ssbtr")(
        m, n, kLocal, int(opA), _scalar_arg(p, a), _ptr(A), lda, _ptr(B), ldb, _scalar_arg(p, b),
        _ptr(C), ldc, cRowOffset, cColOffset, int(cFillMode), distC.handle, ctx.handle))


def pgemm_sbs(mLocal, n, k, alpha, A, lda, B, ldb, bRowOffset, bColOffset,
              distB: MatrixDistribution, beta, C, ldc, ctx: Context, dtype=None):
    """spla::pgemm_sbs (include/spla/pgemm_sbs.hpp:86-118)."""
    p, _ = _prefix(dtype, C, A, B)
    a, b = _scalar(p, alpha), _scalar(p, beta)
    check(getattr(ctx._lib, f"spla_p{p}gemm_sbs")(
        mLocal, n, k, _scalar_arg(p, a), _ptr(A), lda, _ptr(B), ldb, bRowOffset, bColOffset,
        distB.handle, _scalar_arg(p, b), _ptr(C), ldc, ctx.handle))
