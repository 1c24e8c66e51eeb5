/*
 * Extensions of this library beyond SPLA's C API. None of these entry points
 * has a counterpart in SPLA; they are for callers that integrate the engine
 * with their own HIP streams or select the arithmetic of the products.
 */
#ifndef SPLA_AMD_H
#define SPLA_AMD_H

#include "spla/config.h"
#include "spla/context.h"
#include "spla/errors.h"

/* Arithmetic of double and complex-double products. */
enum SplaAmdGemmMode {
  /* Plain FP64 on the FP64 matrix cores (default, SPLA's semantics). */
  SPLA_AMD_GEMM_FP64 = 0,
  /* FP64-accurate split precision (Ozaki scheme II on the INT8 matrix cores)
   * whenever the estimated normwise error bound meets the tolerance; FP64
   * otherwise (non-finite input, too wide an exponent range, too many moduli). */
  SPLA_AMD_GEMM_SPLIT_INT8 = 1,
  /* Split precision only when it is also expected to be faster than FP64. */
  SPLA_AMD_GEMM_AUTO = 2
};

#ifndef __cplusplus
typedef enum SplaAmdGemmMode SplaAmdGemmMode;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Run the context's kernels on the caller's HIP stream (a hipStream_t passed
 * as void*; NULL restores the context's own stream). Calls stay synchronous:
 * they return after the stream has drained. Intended for callers that time or
 * order the library's kernels with their own events (e.g. a torch stream).
 */
SPLA_EXPORT SplaError spla_amd_ctx_set_stream(SplaContext ctx, void* hipStream);

/*
 * Target size in bytes of one k-chunk of host-resident operands streamed to
 * the GPU (default 256 MiB; at least 64 reduction indices per chunk).
 * Smaller chunks start the pipeline earlier, larger ones run longer kernels.
 */
SPLA_EXPORT SplaError spla_amd_ctx_set_stream_chunk_bytes(SplaContext ctx, size_t bytes);

/* Selects the arithmetic of double / complex-double products (see SplaAmdGemmMode). */
SPLA_EXPORT SplaError spla_amd_ctx_set_gemm_mode(SplaContext ctx, SplaAmdGemmMode mode);

/*
 * Normwise error target ||C - C_exact||_F / (||A||_F ||B||_F) of the split
 * precision form (default 1e-14; never looser than the FP64 bound k * 2^-53 of
 * the same product). Must be positive.
 */
SPLA_EXPORT SplaError spla_amd_ctx_set_split_tolerance(SplaContext ctx, double tolerance);

/*
 * Largest device workspace in bytes (residue planes and INT8 GEMM output) of
 * one split-precision product; 0 restores the default, a quarter of the device
 * memory. A product that needs more runs its reduction in chunks of at most
 * this workspace, each chunk a split product of its own added into C in FP64
 * (beta on the first chunk only); one that cannot be chunked (fewer than 4096
 * reduction indices fit) runs on the FP64 kernels (reason 8).
 */
SPLA_EXPORT SplaError spla_amd_ctx_set_split_workspace_limit(SplaContext ctx, size_t bytes);

/*
 * Outcome of the context's last double / complex-double product: used = 1 if
 * it ran in split precision; reason = 0 when used, else why FP64 ran
 * (1 non-finite input, 2 zero operand, 3 exponent range too wide, 4 too many
 * moduli, 5 norm or reduction length out of range, 6 FP64 expected faster, 7 split
 * mode off, 8 split workspace over the limit and not chunkable, or allocation
 * failed; for a product run in reduction chunks: the reason of a chunk that ran on
 * the FP64 kernels);
 * moduli, bitsA and bitsB describe the plan; bound is its estimated normwise
 * error bound. Any output pointer may be NULL.
 */
SPLA_EXPORT SplaError spla_amd_ctx_last_product_info(SplaContext ctx, int* used, int* reason,
                                                     int* moduli, int* bitsA, int* bitsB,
                                                     double* bound);

/* ------------------------------------------------------------------------
 * Batched ring-step block products.
 *
 * One ring step of pgemm_ssb / pgemm_ssbtr forms many independent block
 * products C_b <- alpha op(A_b) B_b + beta C_b with op(A) = A^T or A^H
 * (SPLA src/pgemm_ssb/ring_ssb_gpu.cpp:193-197 issues one vendor GEMM per
 * block); one of pgemm_sbs forms C(:, cols) <- alpha A(:, rows) B_blk +
 * beta' C(:, cols) (src/pgemm_sbs/ring_sbs_gpu.cpp:57-88, 278-282). In FP64
 * mode a batch is one GEMM launch over the tiles and reduction splits of every
 * block plus one fixed-order reduction launch, the two nodes of one hipGraph
 * built at creation and replayed by every execute; tile shape and per-block
 * split counts come from a cost model of the launch or from timed candidates.
 * Split precision runs the blocks one by one; the automatic mode runs the
 * blocks the split form cannot win as one FP64 launch. All pointers must be
 * device memory. A block's fill mode and offsets follow pgemm_ssbtr: tiles
 * wholly outside the requested triangle of the global C are skipped.
 */
typedef void* SplaAmdBatch;

typedef struct SplaAmdBlock {
  int m, n, k;        /* C block m x n, reduction length k */
  const void* A;      /* ssb: k x m, lda >= k; sbs: m x k, lda >= m */
  int lda;
  const void* B;      /* k x n, leading dimension ldb >= k */
  int ldb;
  void* C;            /* m x n, leading dimension ldc >= m */
  int ldc;
  SplaFillMode fillMode;
  int rowOffset;      /* position of the block in the global C */
  int colOffset;
} SplaAmdBlock;

enum SplaAmdElement { SPLA_AMD_DOUBLE = 0, SPLA_AMD_COMPLEX_DOUBLE = 1 };

enum SplaAmdTuning {
  SPLA_AMD_TUNE_MODEL = 0,     /* cost model of the schedule; the context's dispatch */
  SPLA_AMD_TUNE_MEASURE = 1,   /* also time candidate plans (tile shape, split scale,
                                  dispatch) on scratch copies of C */
  SPLA_AMD_TUNE_PERSISTENT = 2 /* cost model, persistent dispatch: one workgroup per
                                  resident slot with a static list of tiles */
};

#ifndef __cplusplus
typedef struct SplaAmdBlock SplaAmdBlock;
typedef enum SplaAmdElement SplaAmdElement;
typedef enum SplaAmdTuning SplaAmdTuning;
#endif

/* Plans (and autotunes) a batch; alpha / beta point to one double or one
 * (re, im) pair. The block array is copied. */
SPLA_EXPORT SplaError spla_amd_batch_create_ssb(SplaAmdBatch* batch, SplaContext ctx,
                                                SplaAmdElement type, SplaOperation opA, int count,
                                                const SplaAmdBlock* blocks, const void* alpha,
                                                const void* beta, SplaAmdTuning tuning);

/* The same for the block products of a pgemm_sbs ring step: op(A) = A with A
 * m x k (lda >= m), B k x n, every block's fill mode SPLA_FILL_MODE_FULL. A
 * ring step applies beta only on the first visit of a column block
 * (ring_sbs_gpu.cpp:278-282): products with beta and with 1 are two batches. */
SPLA_EXPORT SplaError spla_amd_batch_create_sbs(SplaAmdBatch* batch, SplaContext ctx,
                                                SplaAmdElement type, int count,
                                                const SplaAmdBlock* blocks, const void* alpha,
                                                const void* beta, SplaAmdTuning tuning);

/* Runs the batch (synchronous). In split-precision mode every block runs as
 * its own split product. In automatic mode (SPLA_AMD_GEMM_MODE_AUTO) the blocks
 * whose shape rules the split form out join one grouped FP64 launch; the others
 * run split, block by block, unless the sum of their split-form lower bounds
 * does not beat the grouped launch's estimate for them, in which case every
 * block runs in the grouped FP64 launch. */
SPLA_EXPORT SplaError spla_amd_batch_execute(SplaAmdBatch batch);

/* Blocks the last spla_amd_batch_execute ran in the split form, block by block
 * (0: every block in the grouped FP64 launch). No counterpart in the reference,
 * which has no batch object. */
SPLA_EXPORT SplaError spla_amd_batch_split_blocks(SplaAmdBatch batch, int* blocks);

/* Plan summary: tile configuration (0: 128 x 128, 1: 64 x 64 embedded), work
 * items, workgroups of the persistent launch, modelled and measured (0 when
 * not measured) milliseconds. Output pointers may be NULL. */
SPLA_EXPORT SplaError spla_amd_batch_info(SplaAmdBatch batch, int* tileConfig, int* items,
                                          int* workgroups, double* modelMs, double* measuredMs);

/* Dispatch of the batch's GEMM launch: 1 persistent (one workgroup per resident
 * slot working through a static list of tiles), 0 grouped (one workgroup per
 * tile and split). */
SPLA_EXPORT SplaError spla_amd_batch_persistent(SplaAmdBatch batch, int* persistent);

SPLA_EXPORT SplaError spla_amd_batch_destroy(SplaAmdBatch* batch);

/* ------------------------------------------------------------------------
 * Allocation-free product.
 *
 * No counterpart in SPLA: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). spla_zgemm on device-resident A, B and C (the same checks and
 * errors, synchronous), always on the FP64 kernels with the plan spla_zgemm
 * makes in FP64 mode, so both give bit-identical results. The split-K partial
 * sums live in a 64 MiB device array of the library that each GPU holds until
 * the process ends; no device memory is allocated (the first call on a device
 * only looks the array up). Limits: SPLA_GPU_ALLOCATION_ERROR when the plan's
 * partial sums (splits x n x 2m doubles) exceed 64 MiB;
 * SPLA_GPU_INVALID_DEVICE_POINTER_ERROR for host-resident operands. Calls
 * that overlap on one GPU share the array and race.
 */
SPLA_EXPORT SplaError spla_amd_zgemm_noalloc(SplaOperation opA, SplaOperation opB, int m, int n,
                                             int k, const void* alpha, const void* A, int lda,
                                             const void* B, int ldb, const void* beta, void* C,
                                             int ldc, SplaContext ctx);

/* ------------------------------------------------------------------------
 * Ring-step planning (host only, no device work).
 *
 * The blocks a pgemm_ssb / pgemm_ssbtr ring computes and where each piece of
 * a block lives, by SPLA's own rules (block shape
 * src/util/block_size_selection.hpp:57-186 with the constants of
 * src/pgemm_ssb/pgemm_ssb_gpu.cpp:108-112, block order :170-190, fill
 * predicate src/block_generation/block.hpp:58-66, placement
 * block_cyclic_generator.cpp:38-112 and mirror_generator.hpp:40-110). The
 * block products of one ring step are a batch (spla_amd_batch_create_ssb).
 */
typedef struct SplaAmdRingBlock {
  int row, col;          /* inside the m x n sub-matrix of C */
  int numRows, numCols;
} SplaAmdRingBlock;

typedef struct SplaAmdPiece {
  int globalRow, globalCol;  /* first element in the global C */
  int subRow, subCol;        /* relative to the region's origin */
  int localRow, localCol;    /* first element in the owner's local C */
  int numRows, numCols;
  int rank;                  /* owner, column-major over the grid; -1: every rank */
} SplaAmdPiece;

#ifndef __cplusplus
typedef struct SplaAmdRingBlock SplaAmdRingBlock;
typedef struct SplaAmdPiece SplaAmdPiece;
#endif

/* Block shape and ordered active blocks of pgemm_ssb / ssbtr on a distribution
 * of C (type, process grid) over commSize ranks; consecutive groups of
 * commSize blocks form one ring tile. numBlocks receives the total; at most
 * capacity blocks are written (blocks may be NULL). */
SPLA_EXPORT SplaError spla_amd_ssb_ring_blocks(SplaDistributionType distType, int gridRows,
                                               int gridCols, int commSize, int m, int n,
                                               int cRowOffset, int cColOffset,
                                               SplaFillMode cFillMode, int tileSizeHost,
                                               int* rowsInBlock, int* colsInBlock,
                                               int* numBlocks, SplaAmdRingBlock* blocks,
                                               int capacity);

/* Pieces of the region numRows x numCols at (rowOffset, colOffset) of C:
 * block-cyclic blocks of rowsInBlock x colsInBlock on a gridRows x gridCols
 * grid, or (mirror) tiles of that size held by every rank. numPieces receives
 * the total; at most capacity pieces are written (pieces may be NULL). */
SPLA_EXPORT SplaError spla_amd_region_pieces(SplaDistributionType distType, int rowsInBlock,
                                             int colsInBlock, int gridRows, int gridCols,
                                             int numRows, int numCols, int rowOffset,
                                             int colOffset, int* numPieces, SplaAmdPiece* pieces,
                                             int capacity);

/* Kernel phases reported by spla_amd_ctx_last_timing. */
enum SplaAmdPhase {
  SPLA_AMD_PHASE_GEMM_FP64 = 0,        /* FP64 MFMA GEMM kernels */
  SPLA_AMD_PHASE_SPLITK_REDUCE = 1,    /* split-K reduction and alpha / beta update */
  SPLA_AMD_PHASE_SPLIT_STATS = 2,      /* split precision: operand statistics */
  SPLA_AMD_PHASE_SPLIT_RESIDUES = 3,   /* split precision: residue planes */
  SPLA_AMD_PHASE_SPLIT_INT8_GEMM = 4,  /* split precision: batched INT8 GEMM */
  SPLA_AMD_PHASE_SPLIT_CRT = 5,        /* split precision: reconstruction and update */
  SPLA_AMD_NUM_PHASES = 6
};

/*
 * Enables (1) or disables (0) per-phase kernel timing: HIP events are recorded
 * on the launch stream around every phase of the following calls.
 */
SPLA_EXPORT SplaError spla_amd_ctx_set_timing(SplaContext ctx, int enable);

/*
 * Milliseconds spent by each SplaAmdPhase in the context's last call (summed
 * over that call's launches; 0 for phases that did not run). ms must hold
 * SPLA_AMD_NUM_PHASES doubles.
 */
SPLA_EXPORT SplaError spla_amd_ctx_last_timing(SplaContext ctx, double* ms);

/*
 * Process-wide wall-clock timer. SPLA exports these four functions (without a
 * header) when built with SPLA_TIMING (src/timing/timing.cpp:49-86); the
 * Fortran and test code of SPLA call them with (length, pointer) strings.
 * Scopes nest per thread: a scope started while another is open on the same
 * thread becomes its child, and repeated scopes at one place accumulate their
 * durations (seconds). A stop that does not match the innermost open scope is
 * ignored and reported as a warning by spla_timer_print.
 */
SPLA_EXPORT SplaError spla_timer_start(int n, const char* name);
SPLA_EXPORT SplaError spla_timer_stop(int n, const char* name);
/* Writes {"name": {"timings": [...], "sub-timings": {...}}, ...} to the file
 * named by (n, name); SPLA_INVALID_PARAMETER_ERROR if it cannot be opened. */
SPLA_EXPORT SplaError spla_timer_export_json(int n, const char* name);
/* Prints a table (count, total, % of the top-level scope, % of the parent,
 * median, min, max) to standard output. */
SPLA_EXPORT SplaError spla_timer_print(void);

/*
 * Records the library's own scopes ("pgemm_ssb_gpu", "pgemm_sbs_gpu" as in
 * SPLA, "gemm_gpu" around every local product and "split_plan" around the host
 * plan of a split-precision product) into the same timer. Off by default.
 */
SPLA_EXPORT SplaError spla_amd_timer_set_library_scopes(int enable);
/* Drops every recorded timing and warning; scopes still open are discarded. */
SPLA_EXPORT SplaError spla_amd_timer_clear(void);

#ifdef __cplusplus
}
#endif

#endif
