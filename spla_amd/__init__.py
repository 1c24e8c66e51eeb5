"""spla_amd: SPLA's local GEMM engine for AMD MI355X (gfx950).

The product is the C/C++ library ``spla_amd/lib/libspla.so.1`` (SPLA's C ABI
and C++ API over hand-written HIP kernels); this package is its Python mirror.
"""

from spla_amd.api import (  # noqa: F401
    MPI_COMM_SELF,
    MPI_COMM_WORLD,
    Batch,
    Context,
    DistributionType,
    FillMode,
    GemmMode,
    GenericError,
    GPUAllocationError,
    GPUBlasError,
    GPUError,
    GPUInvalidDevicePointerError,
    GPUInvalidValueError,
    GPULaunchError,
    GPUNoDeviceError,
    GPUSupportError,
    InternalError,
    InvalidAllocatorFunctionError,
    InvalidHandleError,
    InvalidParameterError,
    InvalidPointerError,
    MatrixDistribution,
    MPIAllocError,
    MPIError,
    MPIThreadSupportError,
    Operation,
    ProcessingUnit,
    SplaError,
    check,
    gemm,
    pgemm_sbs,
    pgemm_ssb,
    pgemm_ssbtr,
    region_pieces,
    ssb_ring_blocks,
    timer,
)
