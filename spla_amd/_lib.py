"""Loader and ctypes declarations of libspla.so.1 (the C ABI of include/spla/*.h).

torch is imported first when it is installed: torch ships its own HIP runtime
(libamdhip64.so.7) and the library must bind to that same runtime, so that
device pointers of torch tensors are valid for the library. The loader fails
loudly when the library has not been built; there is no fallback.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libspla.so.1"

c_int = ctypes.c_int
c_char = ctypes.c_char
c_double = ctypes.c_double
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
c_u64 = ctypes.c_uint64
P = ctypes.POINTER

# MPICH ABI: MPI_Comm is a 32-bit integer.
MPI_Comm = ctypes.c_int
MPI_COMM_WORLD = 0x44000000
MPI_COMM_SELF = 0x44000001



class SplaAmdBlock(ctypes.Structure):
    """struct SplaAmdBlock of include/spla_amd/spla_amd.h."""
    _fields_ = [("m", c_int), ("n", c_int), ("k", c_int), ("A", c_void_p), ("lda", c_int),
                ("B", c_void_p), ("ldb", c_int), ("C", c_void_p), ("ldc", c_int),
                ("fillMode", c_int), ("rowOffset", c_int), ("colOffset", c_int)]


class SplaAmdRingBlock(ctypes.Structure):
    """struct SplaAmdRingBlock of include/spla_amd/spla_amd.h."""
    _fields_ = [("row", c_int), ("col", c_int), ("numRows", c_int), ("numCols", c_int)]


class SplaAmdPiece(ctypes.Structure):
    """struct SplaAmdPiece of include/spla_amd/spla_amd.h."""
    _fields_ = [(f, c_int) for f in ("globalRow", "globalCol", "subRow", "subCol", "localRow",
                                     "localCol", "numRows", "numCols", "rank")]


_lib = None


def _declare(lib: ctypes.CDLL) -> None:
    def fn(name, *args):
        f = getattr(lib, name)
        f.argtypes = list(args)
        f.restype = c_int
        return f

    fn("spla_ctx_create", P(c_void_p), c_int)
    fn("spla_ctx_destroy", P(c_void_p))
    fn("spla_ctx_processing_unit", c_void_p, P(c_int))
    for name in ("num_tiles", "tile_size_host", "tile_size_gpu", "op_threshold_gpu",
                 "gpu_device_id"):
        fn(f"spla_ctx_{name}", c_void_p, P(c_int))
    for name in ("host", "pinned", "gpu"):
        fn(f"spla_ctx_allocated_memory_{name}", c_void_p, P(c_u64))
    for name in ("num_tiles", "tile_size_host", "tile_size_gpu", "op_threshold_gpu"):
        fn(f"spla_ctx_set_{name}", c_void_p, c_int)
    alloc_t = ctypes.CFUNCTYPE(c_void_p, ctypes.c_size_t)
    free_t = ctypes.CFUNCTYPE(None, c_void_p)
    for name in ("host", "pinned", "gpu"):
        fn(f"spla_ctx_set_alloc_{name}", c_void_p, alloc_t, free_t)

    fn("spla_mat_dis_create_block_cyclic", P(c_void_p), MPI_Comm, c_char, c_int, c_int, c_int,
       c_int)
    fn("spla_mat_dis_create_block_cyclic_fortran", P(c_void_p), c_int, c_char, c_int, c_int,
       c_int, c_int)
    fn("spla_mat_dis_create_blacs_block_cyclic_from_mapping", P(c_void_p), MPI_Comm, P(c_int),
       c_int, c_int, c_int, c_int)
    fn("spla_mat_dis_create_blacs_block_cyclic_from_mapping_fortran", P(c_void_p), c_int,
       P(c_int), c_int, c_int, c_int, c_int)
    fn("spla_mat_dis_create_mirror", P(c_void_p), MPI_Comm)
    fn("spla_mat_dis_create_mirror_fortran", P(c_void_p), c_int)
    fn("spla_mat_dis_destroy", P(c_void_p))
    for name in ("proc_grid_rows", "proc_grid_cols", "row_block_size", "col_block_size", "type"):
        fn(f"spla_mat_dis_{name}", c_void_p, P(c_int))
    fn("spla_mat_dis_comm", c_void_p, P(MPI_Comm))
    fn("spla_mat_dis_comm_fortran", c_void_p, P(c_int))
    fn("spla_mat_dis_set_row_block_size", c_void_p, c_int)
    fn("spla_mat_dis_set_col_block_size", c_void_p, c_int)

    for p, scal in (("s", c_float), ("d", c_double), ("c", c_void_p), ("z", c_void_p)):
        fn(f"spla_{p}gemm", c_int, c_int, c_int, c_int, c_int, scal, c_void_p, c_int, c_void_p,
           c_int, scal, c_void_p, c_int, c_void_p)
        fn(f"spla_p{p}gemm_ssb", c_int, c_int, c_int, c_int, scal, c_void_p, c_int, c_void_p,
           c_int, scal, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p)
        fn(f"spla_p{p}gemm_ssbtr", c_int, c_int, c_int, c_int, scal, c_void_p, c_int, c_void_p,
           c_int, scal, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p)
        fn(f"spla_p{p}gemm_sbs", c_int, c_int, c_int, scal, c_void_p, c_int, c_void_p, c_int,
           c_int, c_int, c_void_p, scal, c_void_p, c_int, c_void_p)

    # include/spla_amd/spla_amd.h
    fn("spla_amd_ctx_set_stream", c_void_p, c_void_p)
    fn("spla_amd_ctx_set_stream_chunk_bytes", c_void_p, ctypes.c_size_t)
    fn("spla_amd_ctx_set_gemm_mode", c_void_p, c_int)
    fn("spla_amd_ctx_set_split_tolerance", c_void_p, c_double)
    fn("spla_amd_ctx_set_split_workspace_limit", c_void_p, ctypes.c_size_t)
    fn("spla_amd_ctx_set_timing", c_void_p, c_int)
    fn("spla_amd_ctx_last_timing", c_void_p, P(c_double))
    fn("spla_timer_start", c_int, ctypes.c_char_p)
    fn("spla_timer_stop", c_int, ctypes.c_char_p)
    fn("spla_timer_export_json", c_int, ctypes.c_char_p)
    fn("spla_timer_print")
    fn("spla_amd_timer_set_library_scopes", c_int)
    fn("spla_amd_timer_clear")
    fn("spla_amd_batch_create_ssb", P(c_void_p), c_void_p, c_int, c_int, c_int, c_void_p,
       c_void_p, c_void_p, c_int)
    fn("spla_amd_batch_create_sbs", P(c_void_p), c_void_p, c_int, c_int, c_void_p, c_void_p,
       c_void_p, c_int)
    fn("spla_amd_batch_persistent", c_void_p, P(c_int))
    fn("spla_amd_batch_split_blocks", c_void_p, P(c_int))
    fn("spla_amd_batch_execute", c_void_p)
    fn("spla_amd_zgemm_noalloc", c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_int,
       c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p)
    fn("spla_amd_batch_info", c_void_p, P(c_int), P(c_int), P(c_int), P(c_double), P(c_double))
    fn("spla_amd_batch_destroy", P(c_void_p))
    fn("spla_amd_ssb_ring_blocks", c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
       c_int, P(c_int), P(c_int), P(c_int), c_void_p, c_int)
    fn("spla_amd_region_pieces", c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
       P(c_int), c_void_p, c_int)
    fn("spla_amd_ctx_last_product_info", c_void_p, P(c_int), P(c_int), P(c_int), P(c_int),
       P(c_int), P(c_double))


def lib() -> ctypes.CDLL:
    """Returns the loaded library (loads it on first use)."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401  (binds the HIP runtime torch ships)
        except ImportError:
            pass
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -m spla_amd.build` (hipcc, gfx950)")
        handle = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)
        _declare(handle)
        _lib = handle
    return _lib
