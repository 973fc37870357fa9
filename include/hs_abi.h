/*
 * hs_abi.h -- C ABI of the MI355X (gfx950) blocked SPD solver library
 * libhs_amd.so.
 *
 * Plain C types only: device pointers are ordinary pointers obtained from
 * hs_malloc (or any hipMalloc'd memory), sizes are int64_t, every entry
 * returns an int status. Matrices must be ordinary (coarse-grained) device
 * memory, not host-mapped or fine-grained allocations: the FP64 trailing
 * update applies its result with device atomics performed in the L2.
 *
 * Storage contract (same as the reference's SymmetricMatrix,
 * src/dataStructures/SymmetricMatrix.hpp:9-35): the lower block triangle of
 * an N x N SPD matrix, split into b x b blocks, n = ceil(N/b) block rows.
 * Block (I,J), I >= J, has ID  J*n - J*(J-1)/2 + (I-J); it is stored row
 * major at element offset ID*b*b. Diagonal blocks hold both triangles.
 * Rows/columns >= N are zero padding. Vectors have n*b entries.
 *
 * Status codes: 0 = ok; -i = argument i invalid (LAPACK style);
 * > 0 from a factorisation = 1-based global index of the first
 * non-positive pivot (the reference produces NaN silently there,
 * MatrixOperations.cpp:24); HS_ERR_* below for runtime failures, with
 * hs_last_error_string() describing the last one on the calling thread.
 *
 * Threading: a handle owns one HIP stream and its scratch workspace. Calls on
 * one handle are asynchronous and ordered by that stream (this replaces the
 * reference's host wait() after every kernel); hs_sync() is the barrier.
 * A handle is not re-entrant: use one handle per host thread.
 *
 * Workspace: a handle's device scratch (partial sums, vectors, inverted
 * diagonal blocks) and its event pool are sized at set-up by
 * hs_workspace_reserve(handle, N, b, M) for the largest system the handle
 * will see; no solver entry allocates. An entry that needs more than was
 * reserved returns HS_ERR_WORKSPACE and does nothing.
 */
#ifndef HS_ABI_H
#define HS_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HS_OK 0
#define HS_ERR_HIP (-1000)
#define HS_ERR_NOMEM (-1001)
#define HS_ERR_UNSUPPORTED (-1002)
#define HS_ERR_NOT_CONVERGED (-1003)
#define HS_ERR_WORKSPACE (-1004)

typedef struct hs_handle_s* hs_handle_t;
typedef struct hs_event_s* hs_event_t;

/* ---- runtime ------------------------------------------------------------
 * Stand-in for the reference's sycl::queue (main.cpp:189-192): a device and
 * an in-order stream. */
int hs_handle_create(hs_handle_t* handle, int device);
int hs_handle_destroy(hs_handle_t handle);
int hs_sync(hs_handle_t handle);
/* Workspace bytes that every entry of this ABI needs at most for a system of
 * order N in blocks of b, with M test points for hs_gp_predict_f64 (0 if
 * unused). Host-only arithmetic. Sizes the reference's per-solver device
 * allocations (CG.cpp:165-234, Cholesky.cpp:24-32) once, up front. */
int hs_workspace_query(int64_t N, int b, int64_t M, size_t* bytes);
/* Sizes the handle's workspace and event pool for (N, b, M); keeps a larger
 * existing reservation. A set-up call: it waits for the handle's streams and
 * may allocate. */
int hs_workspace_reserve(hs_handle_t handle, int64_t N, int b, int64_t M);
/* Bytes currently reserved. */
int hs_workspace_size(hs_handle_t handle, size_t* bytes);
const char* hs_last_error_string(void);
int hs_device_count(int* count);
/* Replaces sycl::malloc_device / free / queue.memcpy (Cholesky.cpp:24-32, CG.cpp:204-229). */
int hs_malloc(void** ptr, size_t bytes);
int hs_free(void* ptr);
/* Pinned host memory; replaces sycl::usm_allocator<.., usm::alloc::host>
 * (SymmetricMatrix.hpp:53, RightHandSide.hpp:29). */
int hs_host_malloc(void** ptr, size_t bytes);
int hs_host_free(void* ptr);
/* Total / free device memory of the handle's device (the reference checks
 * global_mem_size before allocating, CG.cpp:170-202). */
int hs_device_memory(hs_handle_t handle, size_t* free_bytes, size_t* total_bytes);
int hs_memcpy_htod(hs_handle_t handle, void* dst, const void* src, size_t bytes);
int hs_memcpy_dtoh(hs_handle_t handle, void* dst, const void* src, size_t bytes);
int hs_memcpy_dtod(hs_handle_t handle, void* dst, const void* src, size_t bytes);
int hs_memset_zero(hs_handle_t handle, void* dst, size_t bytes);
/* Timing events on the handle's stream (the reference reads SYCL event
 * profiling, CG.cpp:386-397, Cholesky.cpp:292-353). */
int hs_event_create(hs_event_t* ev);
int hs_event_destroy(hs_event_t ev);
int hs_event_record(hs_handle_t handle, hs_event_t ev);
int hs_event_elapsed_ms(hs_event_t start, hs_event_t stop, float* ms);

/* ---- kernel ops: one entry per reference L3 kernel ----------------------
 * Arguments keep the reference's meaning and order after the handle. */

/* POTRF of diagonal block `blockID` (block row `blockRow`), in place, upper
 * triangle zeroed, rows >= N untouched.
 * Replaces MatrixOperations::cholesky_optimizedGPU (MatrixOperations.cpp:124-191)
 * and ::cholesky (:7-62). A non-positive pivot is recorded; read it with
 * hs_get_info. */
int hs_potrf_diag_f64(hs_handle_t h, double* A, int64_t N, int b, int blockID, int blockRow);
/* Sticky first-bad-pivot (1-based global row, 0 = none) since the last reset. */
int hs_get_info(hs_handle_t h, int64_t* info);
int hs_reset_info(hs_handle_t h);

/* B <- B L^{-T} for blockCount blocks of block column blockRow starting at
 * block row blockStart. Replaces MatrixMatrixOperations::triangularSolve_optimizedGPU
 * (MatrixMatrixOperations.cpp:50-107). */
int hs_trsm_panel_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart, int blockCount);

/* D <- D - C C^T (lower triangle) on diagonal blocks blockStart..+blockCount-1.
 * Replaces symmetricMatrixMatrixDiagonal_optimizedGPU (MatrixMatrixOperations.cpp:212-304). */
int hs_syrk_diag_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart, int blockCount, int blockCountXY);

/* A_IJ <- A_IJ - L_Ik L_Jk^T, k = blockRow, I in [blockStart, blockStart+blockCount),
 * blockRow < J < I. Replaces matrixMatrixStep_optimizedGPU3 (MatrixMatrixOperations.cpp:645-778). */
int hs_gemm_trailing_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart, int blockCount, int blockCountXY);

/* Fused trailing update of right-looking step k: SYRK on every diagonal block
 * and GEMM on every off-diagonal block of the trailing triangle (blocks
 * (I,J), k < J <= I < n), FP64 MFMA. */
int hs_syrk_gemm_trailing_f64(hs_handle_t h, double* A, int b, int k, int blockCountXY);

/* result = A x on a sub-range of block rows/columns (reset=0 accumulates).
 * Replaces MatrixVectorOperations::matrixVectorBlock{,_GPU}
 * (MatrixVectorOperations.cpp:8-192). */
int hs_symv_block_f64(hs_handle_t h, const double* A, const double* x, double* result, int b, int blockStart_i,
                      int blockStart_j, int blockCount_i, int blockCount_j, int blockCountXY, int reset);

/* y = A x over the whole packed matrix, reading every stored block once and
 * applying it to both triangles; deterministic (fixed-order) reduction.
 * Replaces matrixVectorBlock_GPU at share 1.0 (CG.cpp:364-369). */
int hs_symv_packed_f64(hs_handle_t h, const double* A, const double* x, double* y, int64_t N, int b);
/* hs_symv_packed_f64 repeated reps times with HIP events around each of its
 * two passes on the handle's stream; pass_ms[0] / pass_ms[1] receive the
 * average durations of the read-once block pass and of the row reduction.
 * Measurement entry; no reference counterpart. */
/* Block rows [0, blockRows) of y = A x (y's other rows untouched): the GPU's
 * share of a mat-vec split by rows between a CPU and a GPU
 * (matrixVectorBlock_GPU over blockStart_i = 0, blockCount_i = blockRows, all
 * columns; CG.cpp:342-398). Reads only the packed prefix of block columns
 * 0 .. blockRows-1. */
int hs_symv_rows_f64(hs_handle_t h, const double* A, const double* x, double* y, int64_t N, int b, int blockRows);
int hs_symv_packed_f64_timed(hs_handle_t h, const double* A, const double* x, double* y, int64_t N, int b, int reps,
                             double* pass_ms);

/* x_J <- L_JJ^{-1} x_J (transposed: L_JJ^{-T} x_J), entries >= N untouched.
 * Replaces MatrixVectorOperations::triangularSolveBlockVector (:282-335). */
int hs_trsv_block_f64(hs_handle_t h, const double* A, double* x, int64_t N, int b, int blockRow, int blockID, int transposed);

/* x_I -= L_{I,blockRow} x_blockRow (transposed: L_{blockRow,I}^T x_blockRow)
 * for I in [blockStart, blockStart+blockCount).
 * Replaces MatrixVectorOperations::matrixVectorColumnUpdate (:337-377). */
int hs_col_update_f64(hs_handle_t h, const double* A, double* x, int b, int blockStart, int blockCount, int blockRow,
                      int blockID, int blockCountXY, int transposed);

/* Vector block ops on entries [blockStart*b, (blockStart+blockCount)*b).
 * Replace VectorOperations::scaleVectorBlock, addVectorBlock, subVectorBlock,
 * scaleAndAddVectorBlock (VectorOperations.cpp:6-82). */
int hs_vec_scale_f64(hs_handle_t h, const double* x, double alpha, double* result, int b, int blockStart, int blockCount);
int hs_vec_add_f64(hs_handle_t h, const double* x, const double* y, double* result, int b, int blockStart, int blockCount);
int hs_vec_sub_f64(hs_handle_t h, const double* x, const double* y, double* result, int b, int blockStart, int blockCount);
int hs_vec_scale_add_f64(hs_handle_t h, const double* x, double alpha, const double* y, double* result, int b,
                         int blockStart, int blockCount);
/* x.y over the block range; fixed-order two-pass reduction; result written
 * to device memory *result_dev (stays on the device, as the CG loop needs).
 * Replaces scalarProduct + sumFinalScalarProduct (VectorOperations.cpp:84-185). */
int hs_dot_f64(hs_handle_t h, const double* x, const double* y, int b, int blockStart, int blockCount, double* result_dev);

/* Streaming-read probe: reads count doubles once (16-byte non-temporal loads,
 * one work-group per 16384 doubles) and writes each work-group's sum to
 * partials_dev (ceil(count / 16384) doubles). Measurement helper that gives
 * the read rate a bandwidth-bound kernel can reach; no reference counterpart. */
int hs_read_sum_f64(hs_handle_t h, const double* x, int64_t count, double* partials_dev);

/* ---- whole-solver paths -------------------------------------------------- */

/* Right-looking blocked Cholesky A -> L in place on the device, stream
 * ordered, no host waits. Returns 0 or the first bad pivot (>0).
 * Replaces Cholesky::solve_heterogeneous at share 1.0 (Cholesky.cpp:413-493). */
int hs_cholesky_f64(hs_handle_t h, double* A, int64_t N, int b);

/* hs_cholesky_f64 with HIP events around every launch: phase_ms[0..2] get
 * the summed durations of the POTRF, TRSM and trailing-update launches and
 * launches[0..2] their counts. Replaces the per-step kernel timing the
 * reference collects into MetricsTracker (Cholesky.cpp:261-359). */
int hs_cholesky_f64_timed(hs_handle_t h, double* A, int64_t N, int b, double* phase_ms, int64_t* launches);

/* The rest-update launches of the look-ahead schedule hs_cholesky_f64 uses
 * for N and b (b a multiple of 64): their number and
 * their algorithmic flops (b^3 per diagonal block and 2 b^3 per off-diagonal
 * block and panel). Measurement helper; no reference counterpart. */
/* hs_cholesky_f64 with a per-column timing record (host arrays of n =
 * ceil(N/b) doubles, ms): potrf_ms[k] the diagonal block k, trsm_ms[k] panel
 * k, update_ms[k] the fused SYRK+GEMM trailing-update launches whose first
 * panel is k (the look-ahead schedule updates several panels per launch, so
 * the other columns of a step hold 0). The per-column vectors of
 * Cholesky::solve_heterogeneous's metrics (Cholesky.cpp:261-359). */
int hs_cholesky_f64_trace(hs_handle_t h, double* A, int64_t N, int b, double* potrf_ms, double* trsm_ms,
                          double* update_ms);
/* The two-stream look-ahead schedule of the factorisations, recorded on the
 * host instead of issued (no device needed): n block columns, `panels` panels
 * per step (0 = the FP64 factor's default, clamped to 1 .. n/2), the mixed
 * factor's bf16 plane splits when planes != 0. Each operation is 8 int64
 * (kind, stream, a0 .. a5): 0 record(event), 1 wait(event), 2 potrf(k),
 * 3 trsm(k, start, count), 4 trail(k, mode, rowStart, rowCount, np,
 * reads_planes), 5 prep(k0); streams 1 (ordering) and 2 (panels); events
 * P[s] = s, U[s] = n + 1 + s. No reference counterpart: the reference's
 * schedule is the sequence with a host wait after each kernel
 * (Cholesky.cpp:163-259). */
int hs_lookahead_dry_run(int n, int panels, int planes, int64_t* ops, int64_t capacity, int64_t* count);
int hs_cholesky_schedule_f64(int64_t N, int b, int64_t* rest_launches, double* rest_flops);

/* The work-group map of one fused trailing-update launch, replayed on the host
 * with the kernels' own decode (no device needed): step k, rows rowStart ..
 * rowStart+rowCount-1, mode 1-6 (1 diagonal blocks, 2 off-diagonal, 3 both,
 * 4 look-ahead rest of the triangle, 5 one look-ahead column, 6 look-ahead
 * columns), np panels, tilesPerBlock work-groups per block pair; band (block
 * rows, 1..16) and order (bit 0 bands, bit 1 one contiguous stretch per XCD)
 * override the library's choice when band > 0 and order >= 0. For work-group
 * g, out[3g .. 3g+2] = (I, J, tile), or -1 three times for the idle
 * work-groups that pad a launch to a multiple of 8; *count = grid size, at
 * most `capacity` work-groups written. No reference counterpart: the
 * reference decodes one (I, J) per work-group with its triangular-number
 * square root (MatrixMatrixOperations.cpp:698-710). */
int hs_trail_map_dry_run(int n, int k, int rowStart, int rowCount, int mode, int np, int tilesPerBlock, int band,
                         int order, int32_t* out, int64_t capacity, int64_t* count);

/* The same for the look-ahead schedule of hs_cholesky_mixed with this variant
 * (b = 128 or 256), whose steps may apply more panels than the FP64 one.
 * Measurement helper; no reference counterpart. */
int hs_cholesky_mixed_schedule(int64_t N, int b, int variant, int64_t* rest_launches, double* rest_flops);

/* Forward L z = x then backward L^T y = z, in place on x (device).
 * Replaces TriangularSystemSolver::solve (TriangularSystemSolver.cpp:15-105). */
int hs_tri_solve_f64(hs_handle_t h, const double* L, double* x, int64_t N, int b);

/* Prepared factor for repeated solves with one L: hs_tri_prepare_f64 writes
 * the inverted diagonal blocks and 1024-row super-blocks of L into `prep`
 * (hs_tri_prepare_query bytes, caller-owned device memory; 0 bytes for block
 * sizes other than 64, 128 and 256, where prep may be NULL), and
 * hs_tri_solve_prepared_f64 then solves L L^T x' = x in place without
 * rebuilding them. Together they replace TriangularSystemSolver::solve
 * (TriangularSystemSolver.cpp:15-105) called many times on one factor. */
int hs_tri_prepare_query(int64_t N, int b, size_t* bytes);
int hs_tri_prepare_f64(hs_handle_t h, const double* L, int64_t N, int b, double* prep);
int hs_tri_solve_prepared_f64(hs_handle_t h, const double* L, const double* prep, double* x, int64_t N, int b);

/* One of the two substitutions, in place on x (device): trans = 0 solves
 * L x' = x, trans = 1 solves L^T x' = x. Replaces the forward or the backward
 * half of TriangularSystemSolver::solve (TriangularSystemSolver.cpp:31-56,
 * 58-82). */
int hs_trsv_packed_f64(hs_handle_t h, const double* L, double* x, int64_t N, int b, int trans);

/* Conjugate gradients with x0 = 0, true residual at iteration % 50 == 0,
 * stop when delta_new <= eps^2 delta_0 or after imax iterations; the loop
 * stays on the device. Outputs may be NULL.
 * Replaces CG::solveHeterogeneous at share 1.0 (CG.cpp:38-163). */
int hs_cg_f64(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b, double eps,
              int64_t imax, int64_t* iterations, double* delta_final, double* delta_zero);

/* hs_cg_f64 with a per-iteration timing record: trace (host, 5 imax doubles)
 * receives for iteration i, in ms, [5i] the mat-vec q = A d, [5i+1] alpha, the
 * x update and the recursive residual update (one fused kernel), [5i+2] the
 * true residual (iterations % 50 == 0, else ~0), [5i+3] delta_new, beta and
 * the d update (fused), [5i+4] the whole iteration. The per-iteration vectors
 * of CG::solveHeterogeneous's metrics (CG.cpp:62-129). */
int hs_cg_f64_trace(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b, double eps,
                    int64_t imax, int64_t* iterations, double* delta_final, double* delta_zero, double* trace);

/* hs_cg_f64 without any allocation and without the handle workspace: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (graph capture, serving loops). Its scratch (SYMV partials, CG
 * vectors, loop state) is a fixed device array of the library,
 * hs_cg_noalloc_limit bytes (32 MiB: N up to about 22000 at b = 128),
 * held by every GPU the library runs on until the process ends; larger
 * systems return HS_ERR_UNSUPPORTED. Calls that overlap on one GPU share that
 * array and race. Same kernels and results as hs_cg_f64. */
int hs_cg_f64_noalloc(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b, double eps,
                      int64_t imax, int64_t* iterations, double* delta_final, double* delta_zero);
int hs_cg_noalloc_limit(size_t* scratch_bytes);

/* The same CG in the reference's own floating-point order, the parity mode:
 * the mat-vec sums every row over block columns 0..n-1 and columns 0..b-1 in
 * order (matrixVectorBlock_GPU, MatrixVectorOperations.cpp:95-192), the dot
 * products are scalarProduct's pair-load and halving tree over work-groups of
 * wg items followed by sumFinalScalarProduct's tree over wgf items
 * (VectorOperations.cpp:84-185; the reference's configuration has wg = 128,
 * wgf = 512, both powers of two up to 1024), the updates are
 * scaleAndAddVectorBlock / subVectorBlock (:45-82), in CG.cpp:38-163's step
 * order. fma = 1 contracts `acc += a * b` and `x + y * a` into fused
 * multiply-adds, fma = 0 rounds every product and sum. 1 <= b <= 256. More
 * than 2 wgf dot partials (N > 4 wg wgf) is the reference's runtime_error:
 * HS_ERR_UNSUPPORTED. Replaces CG::solveHeterogeneous at share 1.0. */
int hs_cg_ref_order_f64(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b, double eps,
                        int64_t imax, int wg, int wgf, int fma, int64_t* iterations, double* delta_final,
                        double* delta_zero);
/* hs_cg_ref_order_f64 with the per-iteration record of hs_cg_f64_trace;
 * [5i+1] is the d.q product, alpha and the x update, [5i+2] the residual
 * (true or recursive), [5i+3] the r.r product, beta and the d update. */
int hs_cg_ref_order_f64_trace(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b,
                              double eps, int64_t imax, int wg, int wgf, int fma, int64_t* iterations,
                              double* delta_final, double* delta_zero, double* trace);
/* The parity mode's mat-vec alone, y = A x over all n block rows
 * (matrixVectorBlock_GPU over the whole matrix, reset = true). */
int hs_symv_ref_order_f64(hs_handle_t h, const double* A, const double* x, double* y, int64_t N, int b, int fma);
/* The parity mode's dot product over blocks [blockStart, blockStart +
 * blockCount) into result_dev[0] (scalarProduct + sumFinalScalarProduct,
 * VectorOperations.cpp:84-185). */
int hs_dot_ref_order_f64(hs_handle_t h, const double* x, const double* y, int b, int blockStart, int blockCount,
                         int wg, int wgf, int fma, double* result_dev);

/* The blocked Cholesky in the reference's own floating-point order, the
 * parity mode of hs_cholesky_f64: per step k the kernels of
 * Cholesky::solve_heterogeneous at share 1.0 (Cholesky.cpp:413-493), each
 * element computed in the sequence of the reference kernel that writes it:
 * POTRF as cholesky_optimizedGPU (MatrixOperations.cpp:124-190), TRSM as
 * triangularSolve_optimizedGPU (MatrixMatrixOperations.cpp:50-107, the
 * multiplication by 1.0 / l_ii), SYRK as symmetricMatrixMatrixDiagonal_
 * optimizedGPU (:212-305, terms subtracted one by one), GEMM as
 * matrixMatrixStep_optimizedGPU3 (:645-779, the reference's default
 * gpuOptimizationLevel 3: terms summed from zero, then subtracted). fma = 1
 * contracts each multiply-subtract / multiply-add statement, fma = 0 rounds
 * every product and sum. b <= 1024, at most 257 block columns. Returns 0 or
 * the first bad pivot (> 0). */
int hs_cholesky_ref_order_f64(hs_handle_t h, double* A, int64_t N, int b, int fma);
/* The forward and backward solves L L^T x = rhs in place on x in the same
 * order (TriangularSystemSolver::solve, TriangularSystemSolver.cpp:15-105:
 * triangularSolveBlockVector, MatrixVectorOperations.cpp:282-335, and
 * matrixVectorColumnUpdate, :337-377), fma as above; entries >= N are left. */
int hs_tri_solve_ref_order_f64(hs_handle_t h, const double* L, double* x, int64_t N, int b, int fma);
/* Its four kernels on the reference's block coordinates (the arguments of
 * the kernels they follow, cited above). */
int hs_potrf_ref_order_f64(hs_handle_t h, double* A, int64_t N, int b, int blockID, int blockRow, int fma);
int hs_trsm_ref_order_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                          int blockCount, int fma);
int hs_syrk_ref_order_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                          int blockCount, int blockCountXY, int fma);
int hs_gemm_ref_order_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                          int blockCount, int blockCountXY, int fma);

/* RBF kernel matrix into the packed layout on the device:
 * K_ij = sf2 * exp(-||x_i - x_j||^2 / (2 ell^2)) + sn2 [i == j],
 * X row-major N x d on the device. Padding entries are written as zero.
 * Replaces MatrixGenerator::generateSPDMatrix's per-block CPU kernel
 * (MatrixGenerator.cpp:107-159) with one device launch. */
int hs_rbf_assemble_f64(hs_handle_t h, const double* X, int64_t N, int d, int b, double ell, double sf2, double sn2,
                        double* A);

/* Reference compatibility features: row i's feature vector is the 8-lag
 * window (t_{i-7}, ..., t_i) of the device series t (zeros before t_0),
 * sf2 = 1, ell = 1, sn2 = 0.01, as MatrixGenerator.cpp:77-99, 125-153. */
int hs_rbf_window_f64(hs_handle_t h, const double* t, int64_t N, int b, double* A);

/* ---- mixed precision (no reference counterpart; the reference's nearest
 * relative is its whole-program FP32 build, CMakeLists.txt:7, 104-106) ---- */

#define HS_MIXED_FP32 0   /* trailing updates on FP32 MFMA */
#define HS_MIXED_BF16X3 1 /* trailing updates on BF16 MFMA, hi+lo split, 3 products */
#define HS_MIXED_F16X3 2  /* trailing updates on FP16 MFMA, hi+lo split of the panels
                             scaled by a power of two (largest entry under 2^15),
                             3 products: ~22-bit operands, near FP32 accuracy */

/* Low-precision factor of the FP64 packed K: L32 (FP32 packed, same layout)
 * <- chol(K + shift I). Diagonal POTRF and panel TRSM in FP64 arithmetic on
 * FP32 storage, trailing SYRK/GEMM per `variant`, FP32 accumulation.
 * matrix_bsz must be 128 or 256. Returns 0 or the first bad pivot (>0). */
int hs_cholesky_mixed(hs_handle_t h, const double* K, float* L32, int64_t N, int b, int variant, double shift);
/* The fused trailing update of right-looking step k on an FP32-stored packed
 * factor (SYRK on every diagonal block, GEMM on every off-diagonal block of
 * the trailing triangle, k < J <= I < n): FP32 MFMA, or BF16 / scaled FP16
 * hi+lo with three products, accumulated in FP32; b % 128 == 0 (the FP16
 * entry scales by the largest |entry| of panel k). The mixed-precision
 * siblings of hs_syrk_gemm_trailing_f64 (symmetricMatrixMatrixDiagonal_optimizedGPU +
 * matrixMatrixStep_optimizedGPU3 of the reference's FP32 build,
 * MatrixMatrixOperations.cpp:212-304, 645-778). */
int hs_syrk_gemm_trailing_f32(hs_handle_t h, float* A, int b, int k, int blockCountXY);
int hs_syrk_gemm_trailing_bf16x3(hs_handle_t h, float* A, int b, int k, int blockCountXY);
int hs_syrk_gemm_trailing_f16x3(hs_handle_t h, float* A, int b, int k, int blockCountXY);
/* Same with per-phase kernel times (as hs_cholesky_f64_timed). */
int hs_cholesky_mixed_timed(hs_handle_t h, const double* K, float* L32, int64_t N, int b, int variant, double shift,
                            double* phase_ms, int64_t* launches);
/* FP64 packed -> FP32 packed copy, shift added to the diagonal. */
int hs_to_f32_packed(hs_handle_t h, const double* K, float* L32, int64_t N, int b, double shift);

typedef struct hs_refine_report_s {
    int64_t ir_iterations;   /* refinement steps taken */
    int64_t pcg_iterations;  /* PCG iterations after a refinement stall (0 if none) */
    int32_t used_pcg;        /* 1 if refinement stalled and PCG finished the solve */
    int32_t converged;       /* 1 if backward_error <= tol */
    double backward_error;   /* ||y - Kx||_inf / (||K||_inf ||x||_inf + ||y||_inf) */
    double relative_residual;/* ||y - Kx||_2 / ||y||_2 */
    double knorm_inf;        /* ||K||_inf */
} hs_refine_report;

/* FP64 iterative refinement of K x = y with the low-precision factor L32:
 * x0 = (L L^T)^{-1} y, then r = y - K x on the FP64 K, x += (L L^T)^{-1} r,
 * until the normwise backward error is <= tol. When a step contracts the
 * backward error by less than 2x, CG preconditioned by L L^T takes over from
 * the current x (at most max_pcg iterations). Returns 0, or
 * HS_ERR_NOT_CONVERGED with the report filled in. */
int hs_refine_solve(hs_handle_t h, const double* K, const float* L32, const double* y, double* x, int64_t N, int b,
                    double tol, int max_ir, int max_pcg, hs_refine_report* report);
/* hs_refine_solve with a second target: the relative residual
 * ||y - K x||_2 / ||y||_2 <= rel_res_tol as well as the backward error
 * <= tol (either one switched off by 0, not both), so that a mixed-precision
 * solve can be run to the residual an FP64 solve reached. */
int hs_refine_solve_to(hs_handle_t h, const double* K, const float* L32, const double* y, double* x, int64_t N, int b,
                       double tol, double rel_res_tol, int max_ir, int max_pcg, hs_refine_report* report);

/* ---- Gaussian-process likelihood (no reference counterpart; the forward
 * pass is the reference's kernel matrix, MatrixGenerator.cpp:143-153, and its
 * Cholesky solve, Cholesky.cpp:413-493, TriangularSystemSolver.cpp:15-105) -- */

/* NLL = 1/2 y^T K^{-1} y + 1/2 log det K + N/2 log(2 pi) for the RBF kernel
 * matrix of X (device, row-major N x d) with hyperparameters ell, sf2, sn2 and
 * targets y (device, N). K (device, packed, n*(n+1)/2*b*b doubles) is
 * overwritten by the factor L. matrix_bsz must be 64, 128 or 256 with
 * matrix_bsz * d <= 4096. Returns 0, or the first bad pivot (>0) when K is
 * not positive definite. */
int hs_gp_nll_f64(hs_handle_t h, const double* X, int64_t N, int d, int b, double ell, double sf2, double sn2,
                  const double* y, double* K, double* nll);

/* hs_gp_nll_f64 plus the gradient grad[0..2] (host) with respect to
 * (log ell, log sf2, log sn2), by reverse-mode differentiation of the blocked
 * Cholesky factorisation and solves. Kbar (device, packed like K) receives
 * dNLL/dK in lower storage: an off-diagonal entry holds the derivative with
 * respect to the (i,j) and (j,i) entries together, upper triangles of
 * diagonal blocks are zero. */
int hs_gp_nll_grad_f64(hs_handle_t h, const double* X, int64_t N, int d, int b, double ell, double sf2, double sn2,
                       const double* y, double* K, double* Kbar, double* nll, double* grad);

/* ---- Gaussian-process prediction (GaussianProcess::start after the solve,
 * GaussianProcess.cpp:42-59) ---- */

/* Test/training kernel matrix on the device, the reference's transposed
 * layout: Kstar[j * N + i] = sf2 * exp(-0.5 / ell^2 * sum_k (Xtrain[i,k] -
 * Xtest[j,k])^2), distance summed in feature order, for i < N training rows
 * and j < M test rows (both row-major with d <= 64 features, on the device).
 * Replaces the CPU-queue kernel of MatrixGenerator::generateTestKernelMatrix
 * (MatrixGenerator.cpp:163-213), which fixes sf2 = ell = 1. */
int hs_rbf_cross_f64(hs_handle_t h, const double* Xtrain, int64_t N, const double* Xtest, int64_t M, int d,
                     double ell, double sf2, double* Kstar);

/* y_r = sum_k A[r * n + k] x_k for r < m (A row-major m x n, device).
 * Replaces MatrixVectorOperations::matrixVectorGP (MatrixVectorOperations.cpp:
 * 379-395); the sum is a fixed-order tree, not the reference's sequential
 * loop. */
int hs_gemv_rows_f64(hs_handle_t h, const double* A, const double* x, double* y, int64_t n, int64_t m);

/* Predictive mean mean_j = sum_i sf2 exp(-0.5/ell^2 ||Xtest_j - Xtrain_i||^2)
 * alpha_i for j < M: hs_rbf_cross_f64 followed by hs_gemv_rows_f64 in one
 * pass that never stores K*. Deterministic (fixed split of the training rows
 * and fixed summation order). alpha (device) holds N values, mean (device) M.
 * Replaces the K* assembly + matrixVectorGP pair of GaussianProcess::start
 * (GaussianProcess.cpp:47-56). */
int hs_gp_predict_f64(hs_handle_t h, const double* Xtrain, int64_t N, const double* Xtest, int64_t M, int d,
                      double ell, double sf2, const double* alpha, double* mean);

/* 8-lag window features of a device series t of N values: X[i][k] =
 * t[i + k - 7], zero before t_0 (X device, row-major N x 8), as
 * MatrixGenerator::readInputVector and the feature loops build them
 * (MatrixGenerator.cpp:77-99, 196-206, 257-272). */
int hs_window_features_f64(hs_handle_t h, const double* t, int64_t N, double* X);

#ifdef __cplusplus
}
#endif

#endif /* HS_ABI_H */
