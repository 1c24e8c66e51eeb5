"""CPU oracle of SPLA's local-GEMM path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this package, and only as the checker or the timed CPU
baseline; the library ``spla_amd/lib/libspla.so.1`` never loads it.

Contents:
  * ctypes bindings of ``oracle.c`` (plain and double-double GEMM
    restatements, SPLA's parameter check) and ``cpu_path.c`` (SPLA's CPU path:
    one host-BLAS ?gemm call);
  * the counter-based input generator (splitmix64) shared by tests and bench;
  * error metrics;
  * the placement restatement lives in ``oracle/placement.py``.
"""

from __future__ import annotations

import ctypes
import glob
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p

_handle = None


def build() -> Path:
    """Compiles the oracle with its Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _handle
    if _handle is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        h = ctypes.CDLL(str(LIB))
        h.oracle_gemm_d.argtypes = [_int, _int, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl,
                                    _vp, _i64]
        h.oracle_gemm_d_dd.argtypes = h.oracle_gemm_d.argtypes
        h.oracle_gemm_z.argtypes = [_int, _int, _i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp,
                                    _vp, _i64]
        h.oracle_gemm_z_dd.argtypes = h.oracle_gemm_z.argtypes
        h.oracle_gemm_z_dd_entries.argtypes = [_int, _int, _i64, _vp, _vp, _i64, _vp, _i64, _vp,
                                               _vp, _i64, _vp, _vp, _i64, _vp]
        h.oracle_gemm_d_dd_entries.argtypes = [_int, _int, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl,
                                               _vp, _i64, _vp, _vp, _i64, _vp]
        h.oracle_check_gemm_param.argtypes = [_int, _int, _i64, _i64, _i64, _vp, _i64, _vp, _i64,
                                              _vp]
        h.oracle_check_gemm_param.restype = _int
        h.cpu_path_init.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                    ctypes.c_char_p]
        h.cpu_path_init.restype = _int
        h.cpu_path_set_threads.argtypes = [_int]
        h.cpu_path_get_threads.restype = _int
        h.cpu_path_dgemm.argtypes = [_int, _int, _int, _int, _int, _dbl, _vp, _int, _vp, _int,
                                     _dbl, _vp, _int]
        h.cpu_path_zgemm.argtypes = [_int, _int, _int, _int, _int, _vp, _vp, _int, _vp, _int, _vp,
                                     _vp, _int]
        _handle = h
    return _handle


def _p(a: np.ndarray | None):
    return None if a is None or a.size == 0 else a.ctypes.data


# ---------------------------------------------------------------------------
# GEMM restatements. Arrays are 1-D numpy buffers holding column-major data
# with the given leading dimensions (exactly the C pointers SPLA receives).

def gemm(opA: int, opB: int, m: int, n: int, k: int, alpha, A, lda, B, ldb, beta, C, ldc,
         accurate: bool = False) -> None:
    """C <- alpha op(A) op(B) + beta C in place (plain or double-double)."""
    h = lib()
    if np.iscomplexobj(C):
        al = np.array([complex(alpha)], dtype=np.complex128)
        be = np.array([complex(beta)], dtype=np.complex128)
        fn = h.oracle_gemm_z_dd if accurate else h.oracle_gemm_z
        fn(opA, opB, m, n, k, _p(al), _p(A), lda, _p(B), ldb, _p(be), _p(C), ldc)
    else:
        fn = h.oracle_gemm_d_dd if accurate else h.oracle_gemm_d
        fn(opA, opB, m, n, k, float(alpha), _p(A), lda, _p(B), ldb, float(beta), _p(C), ldc)


def gemm_entries(opA, opB, k, alpha, A, lda, B, ldb, beta, C0, ldc, rows, cols) -> np.ndarray:
    """Double-double accurate entries (rows[s], cols[s]) of alpha op(A) op(B) + beta C0."""
    h = lib()
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    cnt = len(rows)
    if np.iscomplexobj(A):
        out = np.empty(cnt, dtype=np.complex128)
        al = np.array([complex(alpha)], dtype=np.complex128)
        be = np.array([complex(beta)], dtype=np.complex128)
        h.oracle_gemm_z_dd_entries(opA, opB, k, _p(al), _p(A), lda, _p(B), ldb, _p(be),
                                   _p(C0) if C0 is not None else None, ldc, _p(rows), _p(cols),
                                   cnt, _p(out))
    else:
        out = np.empty(cnt, dtype=np.float64)
        h.oracle_gemm_d_dd_entries(opA, opB, k, float(alpha), _p(A), lda, _p(B), ldb, float(beta),
                                   _p(C0) if C0 is not None else None, ldc, _p(rows), _p(cols),
                                   cnt, _p(out))
    return out


def check_gemm_param(opA, opB, m, n, k, A_present, lda, B_present, ldb, C_present) -> int:
    """SPLA's check_gemm_param as an error code (0, 3 or 4)."""
    one = ctypes.c_double(0.0)
    ptr = ctypes.addressof(one)
    return lib().oracle_check_gemm_param(opA, opB, m, n, k, ptr if A_present else None, lda,
                                         ptr if B_present else None, ldb,
                                         ptr if C_present else None)


# ---------------------------------------------------------------------------
# SPLA's CPU path (host BLAS)

_cpu_path_info = None


def cpu_path() -> dict:
    """Loads host BLAS for the CPU path; returns {'lib', 'vendor'}."""
    global _cpu_path_info
    if _cpu_path_info is not None:
        return _cpu_path_info
    h = lib()
    cands = []
    # MKL's default Intel-OpenMP layer conflicts with the GNU OpenMP runtime that
    # torch loads into the same process (wrong complex results, serial speed);
    # the GNU layer shares that runtime.
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    mkl = "/opt/conda/lib/libmkl_rt.so"
    if os.path.exists(mkl):
        cands.append((mkl, b"", b"MKL_Set_Num_Threads", b"mkl_get_max_threads", "MKL"))
    try:
        import scipy  # noqa: F401
        base = Path(scipy.__file__).resolve().parent.parent / "scipy.libs"
        for f in sorted(glob.glob(str(base / "libscipy_openblas-*.so"))):
            cands.append((f, b"scipy_", b"scipy_openblas_set_num_threads",
                          b"scipy_openblas_get_num_threads", "OpenBLAS"))
    except ImportError:
        pass
    for path, prefix, setter, getter, vendor in cands:
        if h.cpu_path_init(path.encode(), prefix, setter, getter) == 0:
            _cpu_path_info = {"lib": path, "vendor": vendor}
            return _cpu_path_info
    raise RuntimeError("no host BLAS found for SPLA's CPU path")


def cpu_path_set_threads(n: int) -> None:
    cpu_path()
    lib().cpu_path_set_threads(int(n))


def cpu_path_gemm(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) -> None:
    """SPLA's CPU path: gemm_host -> cblas_?gemm (column-major)."""
    cpu_path()
    h = lib()
    if np.iscomplexobj(C):
        al = np.array([complex(alpha)], dtype=np.complex128)
        be = np.array([complex(beta)], dtype=np.complex128)
        h.cpu_path_zgemm(opA, opB, m, n, k, _p(al), _p(A), lda, _p(B), ldb, _p(be), _p(C), ldc)
    else:
        h.cpu_path_dgemm(opA, opB, m, n, k, float(alpha), _p(A), lda, _p(B), ldb, float(beta),
                         _p(C), ldc)


# ---------------------------------------------------------------------------
# counter-based inputs: splitmix64 on (seed, matrix id, linear index)

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_bits(seed: int, matrix_id: int, start: int, count: int) -> np.ndarray:
    """64-bit outputs for draw indices start .. start+count-1."""
    with np.errstate(over="ignore"):
        idx = np.arange(start, start + count, dtype=np.uint64)
        z = (np.uint64(seed) * _GOLDEN + (np.uint64(matrix_id) << np.uint64(40)) + idx
             + _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, matrix_id: int, count: int, complex_: bool = False,
            start: int = 0) -> np.ndarray:
    """Values uniform in [-1, 1) (53-bit); complex: separate re / im draws."""
    draws = count * (2 if complex_ else 1)
    z = splitmix_bits(seed, matrix_id, start * (2 if complex_ else 1), draws)
    u = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    v = 2.0 * u - 1.0
    if complex_:
        return v.view(np.complex128)
    return v


# ---------------------------------------------------------------------------
# error metrics

def normwise_error(C, C_ref, normA: float, normB: float) -> float:
    """||C - C_ref||_F / (||A||_F ||B||_F)."""
    d = np.asarray(C) - np.asarray(C_ref)
    return float(np.linalg.norm(d.ravel()) / (normA * normB)) if normA * normB > 0 else float(
        np.linalg.norm(d.ravel()))
