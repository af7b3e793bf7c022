/* C ABI of the MI355X (gfx950) sparse linear-solve backend.
 *
 * This is the boundary the reference's solver layer would bind for a HIP
 * backend beside DefaultSpace (threaded) and SYCLSpace (SYCL). Paths below are
 * relative to the reference repository; each entry point names the interface
 * it stands in for.
 *
 * Conventions
 *   - Every function returns a status (CIE_HIP_OK = 0) and never throws; the
 *     message of the last failure on the calling thread is cie_hip_last_error().
 *     The reference reports the same conditions as cie::Exception via CIE_CHECK
 *     (cieutils/packages/macros/inc/checks.hpp:17-23).
 *   - Vector arguments are DEVICE pointers (hipMalloc'd, or e.g. torch tensors)
 *     unless a parameter says otherwise. Host arrays are borrowed for the call.
 *   - All work of a context is ordered on its one HIP stream. Functions that
 *     return a host value (inner products, solver statistics) synchronise it.
 *   - One context is not thread-safe; separate contexts are independent.
 *   - A matrix, mask, batch, assembly or multigrid plan belongs to the context
 *     it was created on (its derived forms, workspace and solver scratch live
 *     there). A call that passes a context together with such an object must
 *     pass that context; another one is CIE_HIP_ERROR_INVALID_ARGUMENT.
 */
#ifndef CIE_HIP_H
#define CIE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CIE_HIP_OK 0
#define CIE_HIP_ERROR_INVALID_ARGUMENT 1
#define CIE_HIP_ERROR_SHAPE 2            /* ConjugateGradients_impl.hpp:57-61, CSROperator.cpp:40-44 */
#define CIE_HIP_ERROR_NOT_SPD 3          /* p^T A p <= 0: ConjugateGradients_impl.hpp:145-149 */
#define CIE_HIP_ERROR_BREAKDOWN 4        /* r^T z <= 0 before the loop: ConjugateGradients_impl.hpp:119-131 */
#define CIE_HIP_ERROR_ZERO_DIAGONAL 5    /* DiagonalOperator.cpp:76-78 */
#define CIE_HIP_ERROR_MISSING_DIAGONAL 6 /* DiagonalOperator.cpp:79-82 */
#define CIE_HIP_ERROR_HIP 7
#define CIE_HIP_ERROR_OUT_OF_MEMORY 8
#define CIE_HIP_ERROR_NO_DEVICE 9
#define CIE_HIP_ERROR_NOT_IMPLEMENTED 10 /* NotImplementedException: MaskedJacobiOperator.cpp:64-65 */

/* memory locations for cie_hip_csr_create / cie_hip_memcpy */
#define CIE_HIP_HOST 0
#define CIE_HIP_DEVICE 1

/* value planes of a device CSR matrix */
#define CIE_HIP_PLANE_F64 1
#define CIE_HIP_PLANE_F32 2

typedef struct cie_hip_context_s* cie_hip_context;
typedef struct cie_hip_csr_s* cie_hip_csr;

/* IterativeSolver<TSpace>::Statistics (linalg/packages/solvers/inc/IterativeSolver.hpp:21-25):
 * used both as settings (caps and tolerances) and as results. iterationCount
 * follows ConjugateGradients_impl.hpp:134,184-185: on convergence it is NOT
 * incremented for the final loop trip. */
typedef struct cie_hip_statistics {
    uint64_t iterationCount;
    double absoluteResidual;
    double relativeResidual;
} cie_hip_statistics;

/* Results of the fused solvers beyond Statistics. */
typedef struct cie_hip_solve_info {
    cie_hip_statistics stats;      /* as ConjugateGradients::getStats() */
    int32_t converged;             /* 1 if the tolerance test passed */
    int32_t status;                /* same value the call returns */
    double trueRelativeResidual;   /* ||b - A x|| / ||b - A x0|| in FP64 (if requested) */
    double seconds;                /* device time of the solve (HIP events) */
    uint64_t refinements;          /* mixed: FP64 true-residual refinement steps */
    uint64_t spmvCount;            /* SpMV launches that did work */
    double productSeconds;         /* profiling on: mean duration of the loop's SpMV (q = A p) kernel */
    uint64_t productSamples;       /* profiling on: launches that mean is taken over */
    double initialResidual;        /* ||b - A x0||, the denominator of every relative residual */
    uint64_t reportedIterations;   /* loop trips the reference's solver reports (one row each at verbosity 3,
                                      ConjugateGradients_impl.hpp:175-182): iterationCount + 1 on convergence
                                      in the loop, iterationCount otherwise, 0 when the start check ended the
                                      solve (that report is one Termination row at iteration 0) */
} cie_hip_solve_info;

/* Settings of the mixed-precision refined PCG. */
/* The refinement factor the mixed solvers use by default. Swept over 0.3 .. 1e-4
 * on 3D Q2 and Q4 systems (uniform and lognormal conductivity) and 2D Q4: 0.03
 * was never slower than 0.1 beyond noise and 1.5-3% faster on the large 3D
 * systems (fewer refinement products at equal inner iterations); factors of
 * 0.003 and below cost up to 45% more iterations on 2D Q4. */
#define CIE_HIP_DEFAULT_REFINEMENT_FACTOR 0.03

typedef struct cie_hip_mixed_settings {
    cie_hip_statistics settings;   /* caps and FP64 tolerances, as for cie_hip_pcg_f64 */
    double refinementFactor;       /* refine when ||r_32|| <= factor * ||r|| at the last refinement; a value
                                      outside (0, 1) selects CIE_HIP_DEFAULT_REFINEMENT_FACTOR */
    int32_t restart;               /* 1: reset the search direction at each refinement (classic restarted IR) */
    int32_t flags;                 /* CIE_HIP_MIXED_* bits, 0 for the default form */
} cie_hip_mixed_settings;

/* The inner FP32 product applies the FP32 plane's row-sum correction by
 * default: q_i = (A32 p)_i + c_i p_i with c_i = sum_j (a_ij - fl32(a_ij)),
 * built once per matrix, so the FP32 operator keeps the FP64 row sums (FEM
 * stiffness rows sum to zero away from the boundary, and rounding the values
 * perturbs exactly the smooth modes CG converges last). This flag turns it
 * off: the inner product is then the plain FP32 plane. */
#define CIE_HIP_MIXED_PLAIN_FP32 1

/* ---- errors, devices, contexts ------------------------------------------ */
const char* cie_hip_last_error(void);
const char* cie_hip_status_string(int status);
int cie_hip_device_count(int* count);

/* A context = one device + one stream (+ reduction scratch). stream may be an
 * existing hipStream_t (borrowed, e.g. torch's current stream) or NULL to
 * create one. Replaces the sycl::queue held by SYCLSpace (SYCLSpace.hpp:128). */
int cie_hip_context_create(int device, void* stream, cie_hip_context* context);
/* Destroy the objects created on a context (matrices, masks, batches,
 * assembly and multigrid plans) before the context itself. */
int cie_hip_context_destroy(cie_hip_context context);
int cie_hip_context_synchronize(cie_hip_context context);
int cie_hip_context_stream(cie_hip_context context, void** stream);
/* enable != 0: the fused solvers bracket the first 64 loop SpMV kernels of
 * each solve with HIP events and report their mean duration in
 * cie_hip_solve_info.productSeconds. */
int cie_hip_context_set_profiling(cie_hip_context context, int enable);
/* With profiling on: the duration (HIP events on the context's stream) of the
 * main kernel of the last cie_hip_assemble, cie_hip_element_stiffness or
 * cie_hip_csr_reorder call on this context. */
int cie_hip_context_kernel_seconds(cie_hip_context context, double* seconds);
/* Residual history of the fused solvers (cie_hip_pcg_f64, cie_hip_pcg_mixed) on
 * this context, for the reference's per-iteration report
 * (ConjugateGradients_impl.hpp:175-182 calls IterativeSolver::makeIterationReport
 * after every loop trip). history: device memory of `capacity` doubles, or NULL
 * to stop recording. Loop trip k (reported as iteration k) stores the absolute
 * residual its report shows in history[k], for k < capacity; the relative
 * residual is that over cie_hip_solve_info.initialResidual. In the mixed form
 * it is the FP32 recurrence's ||r||, replaced by the FP64 ||b - A x|| on a
 * trip whose correction was folded into x; an FP32 residual that passes the
 * tolerance triggers the FP64 test on the next trip, so its row can read as
 * converged one row before the last. The buffer is the caller's and must
 * outlive the solves that write it. */
int cie_hip_context_set_residual_history(cie_hip_context context, double* history, int64_t capacity);

/* Device memory for callers without an allocator of their own
 * (cieutils/packages/concurrency/inc/sycl.hpp:35-70 DeviceMemory). */
int cie_hip_malloc(cie_hip_context context, size_t bytes, void** pointer);
int cie_hip_free(cie_hip_context context, void* pointer);
/* kind: 0 host->device, 1 device->host, 2 device->device (SYCLSpace.cpp:316-343). */
int cie_hip_memcpy(cie_hip_context context, void* destination, const void* source, size_t bytes, int kind);

/* ---- CSR matrix (linalg/packages/utilities/inc/CSRView.hpp:11-86) -------- */
/* rowExtents[rows+1], columnIndices[nnz] (sorted per row), entries[nnz].
 * location CIE_HIP_HOST: arrays are uploaded to library-owned device memory.
 * location CIE_HIP_DEVICE: index arrays and FP64 entries are borrowed.
 * planes: CIE_HIP_PLANE_F64 and/or CIE_HIP_PLANE_F32 (FP32 copy of entries). */
int cie_hip_csr_create(cie_hip_context context, int64_t rows, int64_t columns, int64_t nnz,
                       const int32_t* rowExtents, const int32_t* columnIndices, const double* entries,
                       int location, int planes, cie_hip_csr* matrix);
int cie_hip_csr_destroy(cie_hip_csr matrix);
/* Copies the device arrays back to host memory (any pointer may be NULL):
 * rowExtents[rows+1], columnIndices[nnz], FP64 entries[nnz] -- the arrays a
 * CSRView exposes (CSRView.hpp:44-66). */
int cie_hip_csr_copy_out(cie_hip_csr matrix, int32_t* rowExtents, int32_t* columnIndices, double* entries);
int cie_hip_csr_shape(cie_hip_csr matrix, int64_t* rows, int64_t* columns, int64_t* nnz);
/* Rewrite the FP64 entries (and the FP32 plane if present) from device memory;
 * the sparsity pattern is unchanged. */
int cie_hip_csr_update_entries(cie_hip_csr matrix, const double* entries);

/* SpMV launch shape, per value precision (64 or 32). Four kernels:
 *   CIE_HIP_SPMV_STAGED: one wavefront per 64-row tile; coalesced entry
 *     streams, products staged in LDS, each row summed in CSR entry order by
 *     one lane -- rows bit-identical to the sequential CPU loop of
 *     CSROperator.cpp:50-55. width = LDS entries per segment (512, 1024),
 *     unroll = entries per lane in flight (4, 8).
 *   CIE_HIP_SPMV_VECTOR: `width` lanes per row (1..64, power of two) with
 *     strided FMAs and a butterfly sum; unroll 1, 2, 4 or 8. The
 *     SYCLCSROperator layout (SYCLCSROperator.cpp:26-27,71-80) is width = 64.
 *   CIE_HIP_SPMV_TILED: one block of width = 256 threads per tile of whole
 *     rows holding at most 256 * unroll entries (unroll 4, 8 or 16): every
 *     lane streams entries, products are staged in LDS and each row is
 *     summed by a power-of-two group of lanes. Tiles are cut once per matrix
 *     and capacity, on the host, on first use.
 *   CIE_HIP_SPMV_BLOCKS: the tiled kernel over 1 x width blocks of adjacent
 *     columns (width 2, 3 or 4; unroll 8 or 16 blocks per lane): each row's
 *     entries are covered by blocks that start at its next uncovered column,
 *     and a block stores width values, zero where the row has no entry. One
 *     gather loads the width x values of a block. The block form is built on
 *     the device on first use (values refreshed after the entries change);
 *     stored zeros, padding included, contribute nothing to a row sum, also
 *     where x is not finite. Needs at least `width` columns.
 * All are bitwise reproducible run to run for a fixed launch. The default is
 * TILED/256/8; STAGED gives rows bit-identical to the CPU loop. cie_hip_csr_autotune
 * times the candidates of the kernels in `kernels` (a mask of 1 << kernel)
 * on the device and keeps the fastest. */
#define CIE_HIP_SPMV_VECTOR 0
#define CIE_HIP_SPMV_STAGED 1
#define CIE_HIP_SPMV_TILED 2
#define CIE_HIP_SPMV_BLOCKS 3
/*   CIE_HIP_SPMV_SORTED: the BLOCKS kernel (same width and unroll) with each
 *     tile's blocks stored in the order of their first column and a 32-bit key
 *     per block (column offset from the tile's smallest column, position in
 *     the tile): adjacent lanes gather adjacent x values. Products are staged
 *     at the block's position, so results are bitwise those of BLOCKS with the
 *     same width and unroll. Built on the device on first use; needs every
 *     tile's block columns to span fewer than 2^20 columns (else the launch
 *     fails and the autotuner skips it). */
#define CIE_HIP_SPMV_SORTED 4
/*   CIE_HIP_SPMV_PAIRED: rows covered by 1 x wide and 1 x narrow blocks of
 *     adjacent columns (a block is wide where a narrow one would leave
 *     entries of the wide span to a further block), in tiles of whole rows of
 *     at most 256 rows and 256 * unroll blocks of each width, each width's
 *     blocks column-sorted within the tile with the sorted tiles' 32-bit key;
 *     width = 10 * wide + narrow: 53 (1 x 5 and 1 x 3) or 32 (1 x 3 and
 *     1 x 2), unroll 2 or 4. Rows whose entries come in runs of both lengths
 *     (3D Q2 stiffness: runs of 5 and 3 columns) need no padding: 16.6% fewer
 *     FP32 bytes than 1 x 3 blocks on 3D Q2 64^3 (14% in FP64). A matrix with
 *     a row of more than 256 * unroll blocks of one width, or a tile whose
 *     blocks of one width span 2^20 columns or more, does not fit
 *     (CIE_HIP_ERROR_INVALID_ARGUMENT at the first product); the autotuner
 *     skips it there. */
#define CIE_HIP_SPMV_PAIRED 5
/* The tiled kernel reads a narrow copy of the column indices when every run of
 * 4096 consecutive entries spans at most 65535 columns (banded FEM matrices in
 * their natural node order do): 16-bit offsets from each run's smallest
 * column, built on the device on first use. The int32 columns stay the
 * matrix's indices for every other kernel. *bytes receives the bytes per
 * column index the tiled kernel reads: 2 (narrow copy) or 4. The block
 * kernel keeps a narrow copy of its block columns by the same rule. */
int cie_hip_csr_column_bytes(cie_hip_csr matrix, int* bytes);
/* The 1 x width column-block form the CIE_HIP_SPMV_BLOCKS kernel reads (built
 * on first use): *blocks receives its block count (nnz <= blocks * width) and
 * *bytes the bytes per block column it reads (2 or 4, by the rule above). An
 * introspection call with no counterpart in the reference. */
int cie_hip_csr_column_blocks(cie_hip_csr matrix, int width, int64_t* blocks, int* bytes);
/* The block counts of the paired form with width = 10 * wide + narrow (built
 * on first use, with unroll 4 unless a launch of the matrix uses that pair):
 * blocks of each width (the form's bytes per product are
 * (wide * v + 4) * wideBlocks + (narrow * v + 4) * narrowBlocks for v-byte
 * values, plus row extents, x and y). */
int cie_hip_csr_paired_blocks(cie_hip_csr matrix, int width, int64_t* wideBlocks, int64_t* narrowBlocks);
int cie_hip_csr_set_launch(cie_hip_csr matrix, int precision, int kernel, int width, int unroll);
int cie_hip_csr_get_launch(cie_hip_csr matrix, int precision, int* kernel, int* width, int* unroll);
int cie_hip_csr_autotune(cie_hip_csr matrix, int precision, int kernels, int repetitions, double* bestSeconds);

/* LinearOperator<TSpace>::product (linalg/packages/solvers/inc/LinearOperator.hpp:15-33),
 * CSROperator.cpp:33-111 / SYCLCSROperator.cpp:32-89:
 *     out = inScale * out + outScale * A * in
 * inScale == 0 does not read out. Row sums use a fixed order: bitwise
 * reproducible run to run. */
int cie_hip_csr_product_f64(cie_hip_context context, cie_hip_csr matrix, double inScale, const double* in,
                            int64_t inSize, double outScale, double* out, int64_t outSize);
int cie_hip_csr_product_f32(cie_hip_context context, cie_hip_csr matrix, float inScale, const float* in,
                            int64_t inSize, float outScale, float* out, int64_t outSize);

/* ---- LinalgSpaceLike operations (solvers/inc/LinalgSpace.hpp:10-35) ------ */
/* DefaultSpace.cpp:50-187, SYCLSpace.cpp:114-313. innerProduct returns a host
 * value (deterministic two-stage reduction). */
int cie_hip_inner_product_f64(cie_hip_context context, const double* left, const double* right, int64_t size,
                              double* result);
int cie_hip_scale_f64(cie_hip_context context, double* target, const double* source, int64_t size, double scale);
int cie_hip_scale_uniform_f64(cie_hip_context context, double* target, int64_t size, double scale);
int cie_hip_add_f64(cie_hip_context context, double* target, const double* source, int64_t size, double scale);
int cie_hip_assign_f64(cie_hip_context context, double* target, const double* source, int64_t size);
int cie_hip_fill_f64(cie_hip_context context, double* target, int64_t size, double value);
int cie_hip_inner_product_f32(cie_hip_context context, const float* left, const float* right, int64_t size,
                              float* result);
int cie_hip_scale_f32(cie_hip_context context, float* target, const float* source, int64_t size, float scale);
int cie_hip_scale_uniform_f32(cie_hip_context context, float* target, int64_t size, float scale);
int cie_hip_add_f32(cie_hip_context context, float* target, const float* source, int64_t size, float scale);
int cie_hip_assign_f32(cie_hip_context context, float* target, const float* source, int64_t size);
int cie_hip_fill_f32(cie_hip_context context, float* target, int64_t size, float value);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory or wait on the host (code under HIP graph capture, tight serving
 * loops). The deterministic dot product of cie_hip_inner_product_f64, the
 * same two kernels and summation order (bitwise the same value), with its
 * block partials in a static device array of 2048 doubles (16 KiB per device,
 * held until the process ends, its address looked up when a context is
 * created on the device) instead of the context's scratch, and the result
 * written to device memory: *result (device) = left . right. It is
 * asynchronous on the context's stream and calls no allocation, memset or
 * synchronization. Every size fits (the grid is capped at 2048 blocks); a
 * launch that would need more partials is rejected. Calls that overlap on one
 * device share the array and race: serialize them (one stream, or events). */
int cie_hip_inner_product_f64_static(cie_hip_context context, const double* left, const double* right,
                                     int64_t size, double* result);

/* ---- Jacobi preconditioner (solvers/src/DiagonalOperator.cpp) ------------ */
/* makeDiagonalOperator: inverse[i] = 1 / a_ii. Fails like the CPU version
 * (DiagonalOperator.cpp:76-82) on the first row with a zero or missing
 * diagonal; *failedRow (optional, host) receives that row. */
int cie_hip_inverse_diagonal_f64(cie_hip_context context, cie_hip_csr matrix, double* inverse, int64_t* failedRow);
int cie_hip_inverse_diagonal_f32(cie_hip_context context, cie_hip_csr matrix, float* inverse, int64_t* failedRow);
/* The row-sum correction of the FP32 plane that the mixed solvers' inner
 * product adds on the diagonal (no reference counterpart):
 * correction[i] = fl32(sum_j (a_ij - fl32(a_ij))), copied to `correction`
 * (device, rows floats). Needs a square matrix with a diagonal in every row
 * (it is built with the FP32 Jacobi inverse). */
int cie_hip_csr_row_sum_correction(cie_hip_context context, cie_hip_csr matrix, float* correction);
/* DiagonalOperator::product (DiagonalOperator.cpp:37-53): out = inScale*out + outScale*D^-1*in. */
int cie_hip_diagonal_product_f64(cie_hip_context context, const double* inverse, double inScale, const double* in,
                                 double outScale, double* out, int64_t size);
int cie_hip_diagonal_product_f32(cie_hip_context context, const float* inverse, float inScale, const float* in,
                                 float outScale, float* out, int64_t size);

/* ---- Solvers -------------------------------------------------------------- */
/* ConjugateGradients<TSpace>::product(inScale, b, outScale, x) with the Jacobi
 * preconditioner (solvers/impl/ConjugateGradients_impl.hpp:49-201): x is the
 * initial guess; on return x = inScale * x + outScale * solution. Same step
 * order, breakdown checks and convergence test (Comparison_impl.hpp:32-64) as
 * the reference; three fused kernels per iteration, no host round trip inside
 * the loop. trueResidual != 0 also computes info->trueRelativeResidual. */
int cie_hip_pcg_f64(cie_hip_context context, cie_hip_csr matrix, double inScale, const double* b,
                    double outScale, double* x, int64_t size, const cie_hip_statistics* settings,
                    int trueResidual, cie_hip_solve_info* info);

/* Mixed-precision refined PCG (no reference counterpart; the reference only
 * instantiates all-FP32 CG, ConjugateGradients.cpp:10). The inner PCG keeps
 * matrix values, Krylov vectors, its correction and dot products in FP32.
 * Whenever the FP32 residual has fallen by refinementFactor (or looks
 * converged, or the iteration cap is near), the correction is added into the
 * FP64 solution, and the true residual b - A x (FP64 matrix plane) and the
 * convergence test are recomputed in FP64 on the device. The matrix needs
 * both planes. Statistics report the FP64 true residual. */
int cie_hip_pcg_mixed(cie_hip_context context, cie_hip_csr matrix, double inScale, const double* b,
                      double outScale, double* x, int64_t size, const cie_hip_mixed_settings* settings,
                      cie_hip_solve_info* info);

/* IterativeSolver<TS>::makeIterationReport (linalg/packages/solvers/src/
 * IterativeSolver.cpp:17-58): the report text for one solver state. At
 * verbosity >= 2: the table header when status->iterationCount is 0; a row
 * (iteration, absolute and relative residual, each in a 24-bit ANSI colour:
 * green below its tolerance, orange not) for a Termination report, a converged
 * state, or every Iteration report at verbosity >= 3; the closing rule once
 * converged. Below verbosity 2 the text is empty. Writes at most capacity - 1
 * characters and a NUL into out (out may be NULL with capacity 0); *length
 * receives the full length. */
#define CIE_HIP_REPORT_TERMINATION 0
#define CIE_HIP_REPORT_ITERATION 1
int cie_hip_iteration_report(int verbosity, int reportType, const cie_hip_statistics* status,
                             const cie_hip_statistics* settings, char* out, size_t capacity, size_t* length);

/* ---- Batched solves (no reference counterpart) ---------------------------- */
/* Many independent SPD systems of uneven size, stored as the diagonal blocks
 * of one CSR matrix (rows of system s: [offsets[s], offsets[s+1]), columns in
 * the same range; checked at creation). Both value planes are required. Each
 * system runs the reference's Jacobi-PCG (ConjugateGradients_impl.hpp:49-201),
 * all systems advance together, one kernel launch per step for the whole
 * batch, with per-system scalars and convergence masks on the device; one
 * iteration's kernels are captured in a HIP graph (useGraph != 0) and
 * replayed until every system is done. The SpMV walks block tiles (one
 * workgroup per tile of whole rows of one system, at most 2048 entries and
 * rowsPerTile[s] rows: 32, 64, 128 or 256, chosen per system by
 * cie_hip_batch_autotune from timed runs; default 256); the vector steps walk
 * chunks of 1024 rows of one system. After every step the finished systems are
 * dropped from the work lists on the device. The tiles hold matrix entries
 * (CIE_HIP_SPMV_TILED), 1 x 3 column blocks (CIE_HIP_SPMV_BLOCKS, at most 2048
 * per tile), the same blocks in column-sorted tiles (CIE_HIP_SPMV_SORTED,
 * results bitwise those of BLOCKS) or paired 1 x 5 / 1 x 3 tiles
 * (CIE_HIP_SPMV_PAIRED, at most 512 blocks of each width per tile; a matrix
 * with a longer row does not fit): cie_hip_batch_autotune times them where
 * the blocks' padding is small (paired tiles wherever they fit) and keeps the
 * fastest; cie_hip_batch_set_tile_kernel sets the format,
 * cie_hip_batch_set_column_blocks (width 3, or 0 for entries) the unsorted
 * ones. */
typedef struct cie_hip_batch_s* cie_hip_batch;
int cie_hip_batch_create(cie_hip_context context, cie_hip_csr matrix, int32_t systems, const int64_t* offsets,
                         cie_hip_batch* batch);
int cie_hip_batch_destroy(cie_hip_batch batch);
int cie_hip_batch_set_tile_rows(cie_hip_batch batch, const int32_t* rowsPerTile);
int cie_hip_batch_get_tile_rows(cie_hip_batch batch, int32_t* rowsPerTile);
int cie_hip_batch_autotune(cie_hip_batch batch, int precision, int repetitions);
int cie_hip_batch_set_column_blocks(cie_hip_batch batch, int width);
int cie_hip_batch_get_column_blocks(cie_hip_batch batch, int* width);
int cie_hip_batch_set_tile_kernel(cie_hip_batch batch, int kernel);
int cie_hip_batch_get_tile_kernel(cie_hip_batch batch, int* kernel);
/* Solves every system in place: x (device, all rows) is the initial guess and
 * the result. mixed != 0 runs the mixed-precision refined form (settings as
 * for cie_hip_pcg_mixed), else FP64 (settings.settings only). Host outputs
 * (each optional): stats[s] (Statistics as for the single solve; for mixed,
 * the FP64 true residual), status[s] (1 converged, 2 iteration cap, 3 p^T A p
 * <= 0, 4 r^T z <= 0), refinements[s]; summary holds the batch time and the
 * largest iteration count. With the context's profiling on and useGraph 0, the
 * first 8 product kernels are bracketed with HIP events and their mean
 * duration is summary.productSeconds. Returns the first breakdown status, if
 * any. */
int cie_hip_pcg_batched(cie_hip_context context, cie_hip_batch batch, const double* b, double* x,
                        const cie_hip_mixed_settings* settings, int mixed, int useGraph, cie_hip_statistics* stats,
                        int32_t* status, uint64_t* refinements, cie_hip_solve_info* summary);

/* ---- Adjoint gradient (no reference counterpart) -------------------------- */
/* For a stiffness assembled from per-element conductivities,
 *     K(k) = sum_e k_e P_e^T K_e0 P_e
 * (the element integrand is linear in the cell modulus,
 * fem/packages/integrands/impl/LinearIsotropicStiffnessIntegrand_impl.hpp:74,
 * scattered by fem/packages/graph/impl/Assembler_impl.hpp:156-194), and an
 * objective J(u) with K u = f, the adjoint lambda solves K lambda = dJ/du
 * (K is symmetric: cie_hip_pcg_f64 solves it) and
 *     gradient[e] = scale[e] * (-lambda_e^T K_e0 u_e)
 * with scale NULL meaning 1 (dJ/dk), or scale = k for dJ/dg with k = exp(g).
 * dofs holds elements x nodesPerElement row indices, -1 for eliminated
 * (Dirichlet) nodes; K_e0 is nodesPerElement^2, row-major. Device arrays. */
int cie_hip_element_gradient(cie_hip_context context, int64_t elements, int32_t nodesPerElement, const int32_t* dofs,
                             const double* elementMatrix, const double* scale, const double* lambda, const double* u,
                             double* gradient);
/* entries[slots[i][a][b]] += factors[i] * K_e0[a][b] for `count` elements, in
 * list order (slot -1 skipped): the change of K when k_e changes by factors[i].
 * Refreshes the FP32 plane and drops the cached Jacobi diagonal. The matrix
 * must own its FP64 entries. Device arrays. */
int cie_hip_csr_add_element_matrices(cie_hip_csr matrix, int64_t count, int32_t nodesPerElement, const int64_t* slots,
                                     const double* elementMatrix, const double* factors);

/* ---- Parameterised stiffness assembly ------------------------------------ */
/* K = sum_e scale[e] P_e^T K_e0 P_e on the pattern of `matrix`: the element
 * scatter of fem/packages/graph/impl/Assembler_impl.hpp:156-194 for an
 * integrand linear in the cell modulus (LinearIsotropicStiffnessIntegrand_impl.hpp:74),
 * done on the device as a row gather: no atomics, every slot summed in element
 * order from zero, so the values are bitwise those of a sequential scatter in
 * element order (and of cie_sys_lagrange_fill). The plan is made once per
 * pattern: dofs (host) holds elements x nodesPerElement row indices, -1 for
 * eliminated nodes; every kept node pair must be in the pattern
 * (CIE_HIP_ERROR_SHAPE otherwise). */
typedef struct cie_hip_assembly_s* cie_hip_assembly;
int cie_hip_assembly_create(cie_hip_csr matrix, int64_t elements, int32_t nodesPerElement, const int32_t* dofs,
                            cie_hip_assembly* plan);
int cie_hip_assembly_destroy(cie_hip_assembly plan);
/* Overwrites the FP64 entries (and the FP32 plane) of the plan's matrix with
 * sum_e scale[e] K_e[a][b]; scale NULL means 1. perElement 0: elementMatrices
 * is one nodesPerElement^2 matrix K_e0 shared by all elements (row-major);
 * perElement 1: elements x nodesPerElement^2, one matrix per element (the
 * contributions of Assembler::addContribution). Device arrays. Drops the
 * cached Jacobi diagonal. The matrix must own its FP64 entries. */
int cie_hip_assemble(cie_hip_assembly plan, const double* elementMatrices, int perElement, const double* scale);
/* Element stiffness of affine cells (the LinearIsotropicStiffnessIntegrand,
 * K_e = modulus D^T D, fem/packages/integrands/impl/LinearIsotropicStiffnessIntegrand_impl.hpp:74,
 * integrated over the cell as IntegrandProcessor does,
 * fem/packages/utilities/impl/IntegrandProcessor_impl.hpp:519-680): with
 * x = x0 + J_e xi on the reference element and the reference's gradient
 * moments M[i][j] = sum_q w_q grad_i N grad_j N^T (dim^2 x npe^2,
 * cie_sys_lagrange_gradient_moments for tensor-product elements),
 *     out[e] = |det J_e| modulus[e] sum_ij (J_e^-1 J_e^-T)_ij M[i][j],
 * the sums over i then j in order. jacobians: elements x dim x dim,
 * row-major (J[r][c] = dx_r / dxi_c); modulus NULL means 1; dim 2 or 3.
 * poisson2D's cells (ScaleTranslateTransform) are affine. Device arrays;
 * out holds elements x npe^2 and feeds cie_hip_assemble(plan, out, 1, NULL). */
int cie_hip_element_stiffness(cie_hip_context context, int64_t elements, int32_t dim, int32_t nodesPerElement,
                              const double* moments, const double* jacobians, const double* modulus, double* out);
/* out[rows] = sum_e scale[e] P_e^T v_e, elements in increasing order: the
 * vector overload of Assembler::addContribution (Assembler_impl.hpp:197-220).
 * perElement 0: one nodesPerElement vector; 1: elements x nodesPerElement.
 * Device arrays. */
int cie_hip_assemble_vector(cie_hip_assembly plan, const double* elementVectors, int perElement, const double* scale,
                            double* out);

/* ---- Reordering (linalg/packages/utilities/inc/reorder.hpp) -------------- */
/* ReorderingStrategy (reorder.hpp:13-17). */
#define CIE_HIP_REORDER_NONE 0
#define CIE_HIP_REORDER_CUTHILL_MCKEE 1
#define CIE_HIP_REORDER_REVERSE_CUTHILL_MCKEE 2
/* makeReordering (reorder.cpp:23-173), host arrays: map[new] = old row.
 * Breadth-first from the first row of least degree (row length); each row's
 * unvisited columns are queued in column order and that batch is sorted by
 * degree with std::sort; a new component starts at the first unvisited row of
 * least degree; reverse Cuthill-McKee reverses the map. The pattern must be
 * square (columns < rows). NONE, or a pattern without entries, gives the
 * identity. */
int cie_hip_make_reordering(int64_t rows, const int32_t* rowExtents, const int32_t* columnIndices, int strategy,
                            int32_t* map);
/* reorder<..., Reverse> on a CSR matrix (reorder.cpp:215-365), done on the
 * device into a new library-owned matrix: row i of the result is row map[i]
 * of `matrix` (reverse 0), or row j goes to map[j] (reverse 1, the map read as
 * old -> new); columns are renumbered the same way and each row re-sorted.
 * Both value planes follow. map is a host permutation of [0, rows). */
int cie_hip_csr_reorder(cie_hip_csr matrix, const int32_t* map, int reverse, cie_hip_csr* reordered);
/* reorder / reverseReorder of an array (reorder.cpp:368-426) on device
 * vectors, out of place: out[i] = in[map[i]] (reverse 0) or
 * out[map[i]] = in[i] (reverse 1). map is a device array; an index outside
 * [0, size) is CIE_HIP_ERROR_SHAPE. */
int cie_hip_reorder_f64(cie_hip_context context, const int32_t* map, int reverse, const double* in, double* out,
                        int64_t size);
int cie_hip_reorder_f32(cie_hip_context context, const int32_t* map, int reverse, const float* in, float* out,
                        int64_t size);

/* ---- MatrixMarket files (cieutils/packages/io/inc/MatrixMarket.hpp) ------ */
/* MatrixMarket::Output (MatrixMarket.cpp:36-78): host arrays to a file.
 * CSR -> "coordinate real general", one "row column value" line per stored
 * entry in CSR order, 1-based; a vector -> "array real general", size x 1.
 * digits 0 prints values as std::ostream does by default (6 significant
 * digits, the reference's files); digits > 0 prints that many (17 round-trips
 * a double). */
int cie_hip_mm_write_csr_f64(const char* path, int64_t rows, int64_t columns, int64_t nnz, const int32_t* rowExtents,
                             const int32_t* columnIndices, const double* entries, int digits);
int cie_hip_mm_write_csr_f32(const char* path, int64_t rows, int64_t columns, int64_t nnz, const int32_t* rowExtents,
                             const int32_t* columnIndices, const float* entries, int digits);
int cie_hip_mm_write_array_f64(const char* path, const double* values, int64_t size, int digits);
int cie_hip_mm_write_array_f32(const char* path, const float* values, int64_t size, int digits);
/* Reading (no reference counterpart: MatrixMarket::Input is empty,
 * MatrixMarket.hpp:25-26). Coordinate files (real, integer or pattern;
 * general, symmetric or skew-symmetric) become CSR with rows sorted by column,
 * duplicates summed in file order and symmetric mirrors added; general
 * real/integer arrays are kept dense, column-major. The arrays stay valid
 * until cie_hip_mm_free; dense data has no row extents or column indices. */
typedef struct cie_hip_mm_s* cie_hip_mm;
int cie_hip_mm_read(const char* path, cie_hip_mm* data);
int cie_hip_mm_shape(cie_hip_mm data, int64_t* rows, int64_t* columns, int64_t* nnz, int* dense);
int cie_hip_mm_arrays(cie_hip_mm data, const int32_t** rowExtents, const int32_t** columnIndices,
                      const double** entries);
int cie_hip_mm_free(cie_hip_mm data);

/* ---- Masked operators, Jacobi smoothers, p-multigrid (solvers/src) ------- */
/* An ansatz mask, one uint16 per row (fem/packages/graph/src/Assembler.cpp:194-213),
 * uploaded once; the host copy is kept for set-up work. */
typedef struct cie_hip_mask_s* cie_hip_mask;
int cie_hip_mask_create(cie_hip_context context, const uint16_t* values, int64_t size, cie_hip_mask* mask);
int cie_hip_mask_destroy(cie_hip_mask mask);
/* MaskedCSROperator (MaskedCSROperator.cpp:37-118) as a matrix of its own:
 * the entries of rows and columns with mask < threshold, compacted on the
 * device. Its products through cie_hip_csr_product_f64/f32 are the masked
 * products (a masked row sums to 0). Values are copied: a later
 * cie_hip_csr_update_entries on `matrix` does not reach `masked`. */
int cie_hip_csr_masked(cie_hip_csr matrix, cie_hip_mask mask, uint16_t threshold, cie_hip_csr* masked);
/* MaskedIdentityOperator (MaskedIdentityOperator.cpp:14-70, CPU form):
 * out = inScale out + outScale (mask < threshold ? in : 0). */
int cie_hip_masked_identity_f64(cie_hip_context context, cie_hip_mask mask, uint16_t threshold, double inScale,
                                const double* in, double outScale, double* out, int64_t size);
/* JacobiOperator (JacobiOperator.cpp:36-62) over lhs (any matrix handle, a
 * masked one included) with the inverse diagonal `inverseDiagonal` (device;
 * cie_hip_inverse_diagonal_f64 of the full matrix, as the driver's): x = 0,
 * `iterations` times x += (w (in - lhs x)) / d, then out = outScale x + inScale out. */
int cie_hip_jacobi_f64(cie_hip_context context, cie_hip_csr lhs, const double* inverseDiagonal, uint64_t iterations,
                       double relaxation, double inScale, const double* in, double outScale, double* out,
                       int64_t size);
/* MaskedJacobiOperator (MaskedJacobiOperator.cpp:46-161; previous = out before
 * each sweep, the reference's threaded path) and ApproximateSubstitution
 * (ApproximateSubstitution.cpp:27-111). inScale != 0 with more than one sweep
 * is CIE_HIP_ERROR_NOT_IMPLEMENTED, as there; a kept row without a diagonal
 * entry is CIE_HIP_ERROR_MISSING_DIAGONAL. */
int cie_hip_masked_jacobi_f64(cie_hip_context context, cie_hip_csr matrix, cie_hip_mask mask, uint16_t threshold,
                              uint64_t iterations, double relaxation, double inScale, const double* in,
                              double outScale, double* out, int64_t size);
int cie_hip_approximate_substitution_f64(cie_hip_context context, cie_hip_csr matrix, uint64_t iterations,
                                         double inScale, const double* in, double outScale, double* out,
                                         int64_t size);
/* The p-multigrid of poisson2D/solver.cpp:111-232 (solveMultigrid). Grid 0
 * keeps mask < 2 and is Jacobi-PCG to a relative residual of 0.5 (at most 5000
 * steps); grid g = 1..order-1 keeps mask < g + 2 and is 6 Jacobi sweeps with
 * relaxation 2/3 on the masked matrix; a cycle visits the grids from the
 * finest down to 0 and back up from 1, each time adding the grid's update to x
 * and subtracting the masked product of the update from the residual. The
 * residual starts as b (x is expected to hold zeros, as in the driver); cycles
 * run until ||r|| <= relativeTolerance ||b|| (the driver's 1e-6) or maxCycles.
 * The grids are built once per plan. */
typedef struct cie_hip_pmg_s* cie_hip_pmg;
typedef struct cie_hip_pmg_info {
    uint64_t cycles;
    uint64_t coarseIterations;   /* summed iteration counts of the coarse CG solves */
    double initialResidual;      /* ||b|| */
    double residual;             /* the tracked residual norm after the last cycle */
    int32_t converged;
    int32_t status;
    double seconds;              /* device time of the solve */
} cie_hip_pmg_info;
int cie_hip_pmg_create(cie_hip_csr matrix, cie_hip_mask mask, int order, cie_hip_pmg* plan);
int cie_hip_pmg_destroy(cie_hip_pmg plan);
int cie_hip_pmg_solve(cie_hip_pmg plan, const double* b, double* x, int64_t size, double relativeTolerance,
                      uint64_t maxCycles, cie_hip_pmg_info* info);
/* The matrix a plan uses on grid `grid`: 0 <= grid < order gives the matrix
 * masked to mask < grid + 2, compacted to its kept rows and columns in row
 * order (grid 0 is the coarse matrix; the plan's own matrix where the mask
 * keeps every row), grid = -1 the coarse matrix as well (null, for both, when
 * no row has mask < 2).
 * Borrowed: valid while the plan lives; not to be destroyed. Its SpMV launch
 * may be set or autotuned like any matrix's (cie_hip_csr_set_launch,
 * cie_hip_csr_autotune), which changes the grid's sweeps and products only in
 * their summation order. No counterpart in the reference, whose masked
 * operators have no format choice (MaskedCSROperator.cpp:15-118). */
int cie_hip_pmg_grid(cie_hip_pmg plan, int grid, cie_hip_csr* matrix);

#ifdef __cplusplus
}
#endif

#endif /* CIE_HIP_H */
