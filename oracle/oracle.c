/*
 * oracle.c -- CPU FP64 restatement of the reference's blocked SPD solver path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so. No product code path calls into it.
 *
 * Every function restates one reference routine (cited by file:line in
 * TimThuering/Heterogeneous-Solvers) in plain C, keeping that routine's
 * arithmetic order: the reference's share-0 (CPU queue) variants are the
 * ones restated, since those are the reference's CPU path. Work items that the
 * reference runs as independent parallel_for items are spread over OpenMP
 * threads; the order of floating-point operations inside one item is kept.
 *
 * Storage: the reference's packed lower-block-triangle layout. With
 * n = ceil(N/b), block (I,J), I>=J, has ID  J*n - J*(J-1)/2 + (I-J)
 * (SymmetricMatrix.hpp:9-35, MatrixGenerator.cpp:107-120); element (i,j)
 * of a block lives at ID*b*b + i*b + j (row major). Diagonal blocks hold
 * both triangles. Padding rows/columns (N % b != 0) are zero.
 *
 * How each function is pinned (checked independently of this file):
 *  - layout, POTRF, TRSM, SYRK, GEMM, SYMV, TRSV, column update, dot, axpy:
 *    the reference's own golden vectors (tests/golden/reference_goldens.json,
 *    extracted from the reference's GoogleTest files) at 1e-12;
 *  - or_cholesky / or_tri_solve: numpy.linalg.cholesky and scipy cho_solve on
 *    dense copies, and L L^T = A reconstruction;
 *  - or_cg: closed-form solution of a dense solve (numpy) and the iteration
 *    count of a textbook CG written independently in the tests;
 *  - or_dot_tree: the reference's dot-product goldens, exact sums of
 *    integers, and hand-derived results on inputs where the pair-load /
 *    halving-tree order decides the rounding (so that another work-group
 *    size, a sequential sum or an exactly rounded sum each fail);
 *  - or_symv_ordered: bit-for-bit equality with or_symv (fma = 0), which the
 *    mat-vec goldens pin, and a numpy product within the rounding bound
 *    (fma = 1);
 *  - or_cg_gpu_order: the textbook CG of the tests, the closed-form dense
 *    solution, and, at fma = 0, a loop in the tests that composes the pinned
 *    or_symv, or_dot_tree, or_sub and or_scale_add in CG.cpp's step order
 *    (bit for bit);
 *  - or_*_gpu_order (Cholesky): at fma = 0, POTRF, SYRK and GEMM bit for bit
 *    against or_potrf, or_syrk and or_gemm, which the goldens pin; the TRSM
 *    against the reference's triangularSolve_optimizedGPU goldens, and, on
 *    inputs where it decides the rounding, against the multiplication by a
 *    rounded reciprocal (so that a division fails); at fma = 1, single
 *    statements on inputs where contraction changes the rounding, compared
 *    with exactly rounded results (so that a missing or misplaced fma fails);
 *    the whole factor against numpy's Cholesky within the rounding bound;
 *  - or_tri_solve_gpu_order: bit for bit against or_tri_solve at fma = 0;
 *    single statements where contraction changes the rounding at fma = 1;
 *    scipy's cho_solve on the same factor within the rounding bound;
 *  - or_rbf / or_rbf_window: numpy brute force of the textbook RBF formula;
 *  - or_rbf_cross / or_window_features / or_gemv_gp: numpy brute force of the
 *    formula and of the window construction, K* = sf2 exactly where a test
 *    input equals a training input, numpy's dot for the mat-vec, and the GP
 *    identity that the predictive mean at the training inputs equals
 *    y - sn2 alpha for alpha = K^{-1} y;
 *  - synthetic generator: an independent numpy SplitMix64 in the tests.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long long i64;

/* packed ID of block (I,J), I >= J */
i64 or_block_id(int I, int J, int n) {
    return (i64)J * n - (i64)J * (J - 1) / 2 + (I - J);
}

i64 or_packed_elems(i64 N, int b) {
    i64 n = (N + b - 1) / b;
    return n * (n + 1) / 2 * (i64)b * b;
}

/* ---------------------------------------------------------------------------
 * POTRF of one diagonal block, in place.
 * Reference: MatrixOperations::cholesky, src/kernel/MatrixOperations.cpp:7-62.
 * Per column k: column below the diagonal divided by sqrt(a_kk), a_kk set to
 * the root, then the rank-1 update of the lower triangle right of column k,
 * and the upper triangle of the block zeroed. Rows >= N (padding) are left.
 * The reference runs one work item per row with a barrier after the column
 * scale; the sequential loop below keeps that per-element operation order.
 * ------------------------------------------------------------------------- */
void or_potrf(double* A, i64 N, int b, int blockID, int blockRow) {
    double* B = A + (i64)blockID * b * b;
    for (int k = 0; k < b; ++k) {
        const double sqrtDiag = sqrt(B[(i64)k * b + k]);
        for (int i = 0; i < b; ++i) {
            if ((i64)blockRow * b + i < N && i > k) B[(i64)i * b + k] = B[(i64)i * b + k] / sqrtDiag;
        }
        B[(i64)k * b + k] = sqrtDiag;
        for (int i = 0; i < b; ++i) {
            if ((i64)blockRow * b + i >= N) continue;
            for (int j = k + 1; j < b; ++j) {
                if (i >= j) {
                    B[(i64)i * b + j] = B[(i64)i * b + j] - B[(i64)i * b + k] * B[(i64)j * b + k];
                } else {
                    B[(i64)i * b + j] = 0;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Panel TRSM: B_g <- B_g * L^{-T} for blockCount blocks below the diagonal
 * block blockID of block column blockRow.
 * Reference: MatrixMatrixOperations::triangularSolve_optimizedCPU,
 * src/kernel/MatrixMatrixOperations.cpp:109-144 (one item per row of B;
 * forward substitution b_k /= l_kk, then b_j -= l_jk b_k for j > k).
 * ------------------------------------------------------------------------- */
void or_trsm(double* A, int b, int blockID, int blockRow, int blockStart, int blockCount) {
    const double* L = A + (i64)blockID * b * b;
#pragma omp parallel for collapse(2) schedule(static)
    for (int g = 0; g < blockCount; ++g) {
        for (int r = 0; r < b; ++r) {
            double* row = A + ((i64)blockID + (blockStart - blockRow) + g) * b * b + (i64)r * b;
            for (int k = 0; k < b; ++k) {
                const double bk = row[k] / L[(i64)k * b + k];
                row[k] = bk;
                for (int j = k + 1; j < b; ++j) row[j] = row[j] - L[(i64)j * b + k] * bk;
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * SYRK on trailing diagonal blocks: D <- D - C C^T (lower triangle i >= j),
 * D = block (blockStart+g, blockStart+g), C = block (blockStart+g, blockRow).
 * Reference: MatrixMatrixOperations::symmetricMatrixMatrixDiagonal_optimizedCPU,
 * src/kernel/MatrixMatrixOperations.cpp:307-361 (value = D; value -= c_ik c_jk).
 * ------------------------------------------------------------------------- */
void or_syrk(double* A, int b, int blockID, int blockRow, int blockStart, int blockCount, int n) {
#pragma omp parallel for schedule(static)
    for (int g = 0; g < blockCount; ++g) {
        const int diag = blockStart + g;
        double* D = A + or_block_id(diag, diag, n) * b * b;
        const double* C = A + ((i64)blockID + (blockStart - blockRow) + g) * b * b;
        for (int j = 0; j < b; ++j) {
            for (int i = j; i < b; ++i) {
                double value = D[(i64)i * b + j];
                for (int k = 0; k < b; ++k) value = value - C[(i64)i * b + k] * C[(i64)j * b + k];
                D[(i64)i * b + j] = value;
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Trailing GEMM: A_IJ <- A_IJ - L_{I,k} L_{J,k}^T for I in
 * [blockStart, blockStart+blockCount), J in [blockRow+1, I-1], k = blockRow.
 * Reference: MatrixMatrixOperations::matrixMatrixStep_optimizedCPU2,
 * src/kernel/MatrixMatrixOperations.cpp:845-906 (sum over k first, then
 * subtract). The block enumeration (row by row of the trailing triangle,
 * skipping rows above blockStart) is the same set of (I,J) pairs.
 * ------------------------------------------------------------------------- */
void or_gemm(double* A, int b, int blockID, int blockRow, int blockStart, int blockCount, int n) {
    /* flatten (I,J) pairs for the parallel loop */
    int pairs = 0;
    for (int I = blockStart; I < blockStart + blockCount; ++I) pairs += I - (blockRow + 1);
    if (pairs <= 0) return;
    int* PI = (int*)malloc(sizeof(int) * pairs);
    int* PJ = (int*)malloc(sizeof(int) * pairs);
    int p = 0;
    for (int I = blockStart; I < blockStart + blockCount; ++I)
        for (int J = blockRow + 1; J < I; ++J) { PI[p] = I; PJ[p] = J; ++p; }
#pragma omp parallel for collapse(2) schedule(static)
    for (int q = 0; q < pairs; ++q) {
        for (int i = 0; i < b; ++i) {
            const int I = PI[q], J = PJ[q];
            double* C = A + or_block_id(I, J, n) * b * b;
            const double* Bm = A + ((i64)blockID + (I - blockRow)) * b * b;
            const double* Cm = A + ((i64)blockID + (J - blockRow)) * b * b;
            for (int j = 0; j < b; ++j) {
                double value = 0;
                for (int k = 0; k < b; ++k) value += Bm[(i64)j * b + k] * Cm[(i64)i * b + k];
                C[(i64)j * b + i] -= value;
            }
        }
    }
    free(PI);
    free(PJ);
}

/* ---------------------------------------------------------------------------
 * Symmetric packed mat-vec on a sub-range of block rows/columns.
 * Reference: MatrixVectorOperations::matrixVectorBlock_CPU,
 * src/kernel/MatrixVectorOperations.cpp:194-280 (and the identical arithmetic
 * of matrixVectorBlock :8-93): per row, blocks J <= I read row-wise, blocks
 * J > I read from the stored block (J,I) transposed; reset=0 accumulates.
 * ------------------------------------------------------------------------- */
void or_symv(const double* A, const double* x, double* y, int b, int bsi, int bsj, int bci, int bcj, int n, int reset) {
#pragma omp parallel for schedule(static)
    for (int li = 0; li < bci * b; ++li) {
        const int i = li + bsi * b;
        const int iIn = i % b;
        const int bi = i / b;
        double acc = 0;
        if (!reset) acc += y[i];
        int bj = bsj;
        const int last = bsj + bcj - 1;
        for (; bj <= (bi < last ? bi : last); ++bj) {
            const double* row = A + or_block_id(bi, bj, n) * b * b + (i64)iIn * b;
            for (int j = 0; j < b; ++j) acc += row[j] * x[(i64)bj * b + j];
        }
        for (; bj < bsj + bcj; ++bj) {
            const double* blk = A + or_block_id(bj, bi, n) * b * b;
            for (int j = 0; j < b; ++j) acc += blk[(i64)j * b + iIn] * x[(i64)bj * b + j];
        }
        y[i] = acc;
    }
}

/* ---------------------------------------------------------------------------
 * TRSV on one diagonal block: x_J <- L_JJ^{-1} x_J (or L_JJ^{-T} x_J).
 * Reference: MatrixVectorOperations::triangularSolveBlockVector,
 * src/kernel/MatrixVectorOperations.cpp:282-335 (entries k >= N untouched).
 * ------------------------------------------------------------------------- */
void or_trsv_block(const double* A, double* x, i64 N, int b, int blockRow, int blockID, int transposed) {
    const double* L = A + (i64)blockID * b * b;
    double* xb = x + (i64)blockRow * b;
    for (int s = 0; s < b; ++s) {
        const int k = transposed ? b - (s + 1) : s;
        const double bk = xb[k] / L[(i64)k * b + k];
        if ((i64)blockRow * b + k >= N) continue;
        xb[k] = bk;
        for (int i = 0; i < b; ++i) {
            const int cond = transposed ? (i < k) : (i > k);
            if (!cond) continue;
            if (!transposed) xb[i] = xb[i] - L[(i64)i * b + k] * bk;
            else xb[i] = xb[i] - L[(i64)k * b + i] * bk;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Column update of the triangular solves: x_I -= L_{I,blockRow} x_blockRow
 * (or x_I -= L_{blockRow,I}^T x_blockRow when transposed) for I in
 * [blockStart, blockStart+blockCount).
 * Reference: MatrixVectorOperations::matrixVectorColumnUpdate,
 * src/kernel/MatrixVectorOperations.cpp:337-377 (sum first, then subtract).
 * ------------------------------------------------------------------------- */
void or_col_update(const double* A, double* x, int b, int blockStart, int blockCount, int blockRow, int blockID, int n, int transposed) {
    const i64 total = (i64)n * (n + 1) / 2;
#pragma omp parallel for schedule(static)
    for (int g = 0; g < blockCount; ++g) {
        i64 id;
        if (!transposed) id = (i64)blockID + (blockStart - blockRow + g);
        else {
            const i64 m = n - g - blockStart;
            id = total - m * (m + 1) / 2 + blockRow - g - blockStart;
        }
        const double* M = A + id * b * b;
        const double* x0 = x + (i64)blockRow * b;
        double* xi = x + (i64)(blockStart + g) * b;
        for (int li = 0; li < b; ++li) {
            double sum = 0.0;
            if (!transposed) for (int k = 0; k < b; ++k) sum += M[(i64)li * b + k] * x0[k];
            else for (int k = 0; k < b; ++k) sum += M[(i64)k * b + li] * x0[k];
            xi[li] -= sum;
        }
    }
}

/* ---------------------------------------------------------------------------
 * Dot product over block range, CPU variant.
 * Reference: VectorOperations::scalarProduct_CPU + sumFinalScalarProduct_CPU,
 * src/kernel/VectorOperations.cpp:187-238: `chunks` contiguous chunks of
 * ceil(len/chunks) elements, each summed in order, then the chunk sums summed
 * in order (the reference uses the CPU's compute-unit count for `chunks`).
 * ------------------------------------------------------------------------- */
double or_dot(const double* x, const double* y, int b, int blockStart, int blockCount, int chunks) {
    const i64 len = (i64)blockCount * b;
    const i64 per = (len + chunks - 1) / chunks;
    double* part = (double*)calloc(chunks, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int c = 0; c < chunks; ++c) {
        const i64 s = (i64)c * per + (i64)blockStart * b;
        i64 e = s + per;
        if (e > (i64)blockStart * b + len) e = (i64)blockStart * b + len;
        double r = 0.0;
        for (i64 i = s; i < e; ++i) r += x[i] * y[i];
        part[c] = r;
    }
    double r = 0.0;
    for (int c = 0; c < chunks; ++c) r += part[c];
    free(part);
    return r;
}

/* Vector block ops. Reference: VectorOperations.cpp:6-82
 * (scaleVectorBlock :6-23 r=x*a, addVectorBlock :25-43 r=x+y,
 *  subVectorBlock :45-62 r=x-y, scaleAndAddVectorBlock :64-82 r=x+y*a). */
void or_scale(const double* x, double a, double* r, int b, int bs, int bc) {
    for (i64 i = (i64)bs * b; i < (i64)(bs + bc) * b; ++i) r[i] = x[i] * a;
}
void or_add(const double* x, const double* y, double* r, int b, int bs, int bc) {
    for (i64 i = (i64)bs * b; i < (i64)(bs + bc) * b; ++i) r[i] = x[i] + y[i];
}
void or_sub(const double* x, const double* y, double* r, int b, int bs, int bc) {
    for (i64 i = (i64)bs * b; i < (i64)(bs + bc) * b; ++i) r[i] = x[i] - y[i];
}
void or_scale_add(const double* x, double a, const double* y, double* r, int b, int bs, int bc) {
    for (i64 i = (i64)bs * b; i < (i64)(bs + bc) * b; ++i) r[i] = x[i] + y[i] * a;
}

/* ---------------------------------------------------------------------------
 * Whole blocked Cholesky, in place (A -> L), CPU path order.
 * Reference: Cholesky::solve_heterogeneous, src/solver/cholesky/Cholesky.cpp:
 * 413-493 at share 0 (:162-172 POTRF `cholesky`, :189-190 TRSM, :212-213
 * SYRK, :224-233 GEMM with cpu_opt=2, rows k+2 .. n-1).
 * ------------------------------------------------------------------------- */
void or_cholesky(double* A, i64 N, int b) {
    const int n = (int)((N + b - 1) / b);
    for (int k = 0; k < n; ++k) {
        const int blockID = (int)or_block_id(k, k, n);
        or_potrf(A, N, b, blockID, k);
        const int cnt = n - (k + 1);
        if (cnt > 0) {
            or_trsm(A, b, blockID, k, k + 1, cnt);
            or_syrk(A, b, blockID, k, k + 1, cnt, n);
        }
        if (cnt > 1) or_gemm(A, b, blockID, k, k + 2, cnt - 1, n);
    }
}

/* ---------------------------------------------------------------------------
 * Forward L z = x then backward L^T y = z, in place on x.
 * Reference: TriangularSystemSolver::solve,
 * src/solver/cholesky/TriangularSystemSolver.cpp:15-105.
 * ------------------------------------------------------------------------- */
void or_tri_solve(const double* L, double* x, i64 N, int b) {
    const int n = (int)((N + b - 1) / b);
    for (int j = 0; j < n; ++j) {
        const int id = (int)or_block_id(j, j, n);
        or_trsv_block(L, x, N, b, j, id, 0);
        if (j < n - 1) or_col_update(L, x, b, j + 1, n - (j + 1), j, id, n, 0);
    }
    for (int j = n - 1; j >= 0; --j) {
        const int id = (int)or_block_id(j, j, n);
        or_trsv_block(L, x, N, b, j, id, 1);
        if (j > 0) or_col_update(L, x, b, 0, j, j, id, n, 1);
    }
}

/* ---------------------------------------------------------------------------
 * Conjugate gradients, share 0 (all rows on the CPU queue).
 * Reference: CG::solveHeterogeneous, src/solver/cg/CG.cpp:38-163, with
 * initCG :277-340 (r = b - A x, d = r, delta0 = delta = r.r), compute_q
 * :342-398, compute_alpha :400-431, update_x :433-442, the true residual at
 * iteration % 50 == 0 :95-101 and :444-473, update_r :475-484,
 * compute_delta_new :486-517, compute_d :519-528, and the stop rule
 * delta_new > eps^2 delta_0 (:62, :78). x starts at zero (the reference's
 * host x is zero; its device copy is never initialised, CG.cpp:217).
 * Returns the iteration count; *delta_out receives the final delta_new.
 * ------------------------------------------------------------------------- */
i64 or_cg(const double* A, const double* rhs, double* x, i64 N, int b, double eps, i64 imax, int chunks, double* delta_out) {
    const int n = (int)((N + b - 1) / b);
    const i64 len = (i64)n * b;
    double* r = (double*)calloc(len, sizeof(double));
    double* d = (double*)calloc(len, sizeof(double));
    double* q = (double*)calloc(len, sizeof(double));
    memset(x, 0, sizeof(double) * len);
    or_symv(A, x, r, b, 0, 0, n, n, n, 1);
    or_sub(rhs, r, r, b, 0, n);
    memcpy(d, r, sizeof(double) * len);
    double delta_new = or_dot(r, r, b, 0, n, chunks);
    const double delta0 = delta_new;
    const double eps2 = eps * eps;
    i64 it = 0;
    while (it < imax && delta_new > eps2 * delta0) {
        or_symv(A, d, q, b, 0, 0, n, n, n, 1);
        const double dq = or_dot(d, q, b, 0, n, chunks);
        const double alpha = delta_new / dq;
        or_scale_add(x, alpha, d, x, b, 0, n);
        if (it % 50 == 0) {
            or_symv(A, x, r, b, 0, 0, n, n, n, 1);
            or_sub(rhs, r, r, b, 0, n);
        } else {
            or_scale_add(r, -1.0 * alpha, q, r, b, 0, n);
        }
        const double delta_old = delta_new;
        delta_new = or_dot(r, r, b, 0, n, chunks);
        const double beta = delta_new / delta_old;
        or_scale_add(r, beta, d, d, b, 0, n);
        ++it;
    }
    if (delta_out) *delta_out = delta_new;
    free(r);
    free(d);
    free(q);
    return it;
}

/* ---------------------------------------------------------------------------
 * The reference's share-1.0 GPU CG path, in its own arithmetic order.
 *
 * At share 1.0 the reference runs CG entirely on the GPU queue
 * (src/solver/cg/CG.cpp:38-163 with blockCountGPU = n) and its kernels fix
 * the order of every floating-point operation:
 *  - q = A d and the true residual use matrixVectorBlock_GPU
 *    (src/kernel/MatrixVectorOperations.cpp:95-192): one work item per row,
 *    resultValue = 0 then += over block columns J = 0..n-1 in order and
 *    columns j = 0..b-1 inside each block, the stored block read row-wise
 *    for J <= I and block (J, I) read transposed for J > I. initCG uses
 *    matrixVectorBlock (:8-93), which has the same per-row order.
 *  - dot products use scalarProduct (src/kernel/VectorOperations.cpp:84-139):
 *    a work-group of wg = conf::workGroupSizeVector (128) items covers 2 wg
 *    consecutive elements; item l starts from 0, adds x[l] y[l], then
 *    x[l + wg] y[l + wg], and the work-group sums its wg values with the
 *    halving tree cache[l] += cache[l + s], s = wg/2 .. 1. Then
 *    sumFinalScalarProduct (:141-185) does the same with
 *    wgf = conf::workGroupSizeFinalScalarProduct (512) items over the
 *    partials (and throws when there are more than 2 wgf of them).
 *  - updates use scaleAndAddVectorBlock (x + y a, :64-82) and subVectorBlock
 *    (x - y, :45-62); r = r - alpha q is scaleAndAdd with -1.0 * alpha.
 * The only freedom left is whether the compiler contracts `acc += a * b`
 * and `x + y * a` into one fused multiply-add; the reference's CMakeLists.txt
 * sets no floating-point flag. fma = 0 evaluates every product and sum
 * separately (this file is built with -ffp-contract=off); fma = 1 contracts
 * exactly those three statement forms (the matrix-vector accumulation, the
 * dot product's second pair term, the scale-and-add) with fma().
 * ------------------------------------------------------------------------- */
static inline double mac(double a, double x, double acc, int fma_mode) {
    return fma_mode ? fma(a, x, acc) : acc + a * x;
}

double or_dot_tree(const double* x, const double* y, int b, int blockStart, int blockCount, int wg, int wgf,
                   int fma_mode) {
    const i64 len = (i64)blockCount * b;
    const i64 off = (i64)blockStart * b;
    const i64 half = len / 2;
    const i64 groups = (half + wg - 1) / wg;
    if (groups > 2 * (i64)wgf) return NAN; /* the reference throws here */
    double* part = (double*)calloc(groups > 0 ? groups : 1, sizeof(double));
    double* cache = (double*)malloc(sizeof(double) * (wg > wgf ? wg : wgf));
    for (i64 g = 0; g < groups; ++g) {
        for (int l = 0; l < wg; ++l) {
            const i64 gi = off + g * 2 * wg + l;
            double c = 0.0;
            if (gi < off + len) c = mac(x[gi], y[gi], c, fma_mode);
            if (gi + wg < off + len) c = mac(x[gi + wg], y[gi + wg], c, fma_mode);
            cache[l] = c;
        }
        for (int st = wg / 2; st > 0; st /= 2)
            for (int l = 0; l < st; ++l) cache[l] = cache[l] + cache[l + st];
        part[g] = cache[0];
    }
    for (int l = 0; l < wgf; ++l) {
        double c = 0.0;
        if (l < groups) c = c + part[l];
        if (l + wgf < groups) c = c + part[l + wgf];
        cache[l] = c;
    }
    for (int st = wgf / 2; st > 0; st /= 2)
        for (int l = 0; l < st; ++l) cache[l] = cache[l] + cache[l + st];
    const double r = cache[0];
    free(cache);
    free(part);
    return r;
}

/* y = A x over all n block rows, each row summed in matrixVectorBlock_GPU's
 * order. The loops run block row by block row with the b row sums of a block
 * row kept side by side, so that every stored block is streamed once per
 * block row that needs it; each row's sequence of operations is exactly the
 * per-row order above (pinned bit for bit against or_symv by the tests). */
void or_symv_ordered(const double* A, const double* x, double* y, int b, int n, int fma_mode) {
#pragma omp parallel for schedule(dynamic)
    for (int I = 0; I < n; ++I) {
        double* acc = y + (i64)I * b;
        for (int i = 0; i < b; ++i) acc[i] = 0.0;
        for (int J = 0; J <= I; ++J) {
            const double* blk = A + or_block_id(I, J, n) * b * b;
            const double* xJ = x + (i64)J * b;
            for (int i = 0; i < b; ++i) {
                double a = acc[i];
                const double* row = blk + (i64)i * b;
                for (int j = 0; j < b; ++j) a = mac(row[j], xJ[j], a, fma_mode);
                acc[i] = a;
            }
        }
        for (int J = I + 1; J < n; ++J) {
            const double* blk = A + or_block_id(J, I, n) * b * b;
            const double* xJ = x + (i64)J * b;
            for (int j = 0; j < b; ++j) {
                const double xj = xJ[j];
                const double* col = blk + (i64)j * b;
                for (int i = 0; i < b; ++i) acc[i] = mac(col[i], xj, acc[i], fma_mode);
            }
        }
    }
}

static void axpy_ordered(const double* x, double a, const double* y, double* r, i64 len, int fma_mode) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < len; ++i) r[i] = fma_mode ? fma(y[i], a, x[i]) : x[i] + y[i] * a;
}

/* CG at share 1.0 on the GPU queue, CG.cpp:38-163 with initCG :277-340,
 * compute_q :342-398 (gpuOptimizationLevel >= 1: matrixVectorBlock_GPU),
 * compute_alpha :400-431, update_x :433-442, computeRealResidual at
 * iteration % 50 == 0 :444-473, update_r :475-484, compute_delta_new
 * :486-517, compute_d :519-528; stop on delta_new > eps^2 delta_0 failing
 * (:62, :78); x starts at zero. wg / wgf are the two work-group sizes of the
 * dot product (128 and 512 in the reference's configuration). Returns the
 * iteration count, or -1 when the dot product's partial count exceeds 2 wgf
 * (the reference's runtime_error). */
i64 or_cg_gpu_order(const double* A, const double* rhs, double* x, i64 N, int b, double eps, i64 imax, int wg,
                    int wgf, int fma_mode, double* delta_out) {
    const int n = (int)((N + b - 1) / b);
    const i64 len = (i64)n * b;
    double* r = (double*)calloc(len, sizeof(double));
    double* d = (double*)calloc(len, sizeof(double));
    double* q = (double*)calloc(len, sizeof(double));
    memset(x, 0, sizeof(double) * len);
    or_symv_ordered(A, x, r, b, n, fma_mode);
    or_sub(rhs, r, r, b, 0, n);
    memcpy(d, r, sizeof(double) * len);
    double delta_new = or_dot_tree(r, r, b, 0, n, wg, wgf, fma_mode);
    i64 it = 0;
    if (isnan(delta_new)) it = -1;
    const double delta0 = delta_new;
    const double eps2 = eps * eps;
    while (it >= 0 && it < imax && delta_new > eps2 * delta0) {
        or_symv_ordered(A, d, q, b, n, fma_mode);
        const double dq = or_dot_tree(d, q, b, 0, n, wg, wgf, fma_mode);
        const double alpha = delta_new / dq;
        axpy_ordered(x, alpha, d, x, len, fma_mode);
        if (it % 50 == 0) {
            or_symv_ordered(A, x, r, b, n, fma_mode);
            or_sub(rhs, r, r, b, 0, n);
        } else {
            axpy_ordered(r, -1.0 * alpha, q, r, len, fma_mode);
        }
        const double delta_old = delta_new;
        delta_new = or_dot_tree(r, r, b, 0, n, wg, wgf, fma_mode);
        const double beta = delta_new / delta_old;
        axpy_ordered(r, beta, d, d, len, fma_mode);
        ++it;
    }
    if (delta_out) *delta_out = delta_new;
    free(r);
    free(d);
    free(q);
    return it;
}

/* ---------------------------------------------------------------------------
 * The reference's share-1.0 GPU blocked Cholesky, in its arithmetic order.
 * Cholesky::solve_heterogeneous at share 1.0 (Cholesky.cpp:413-493) runs per
 * step k: MatrixOperations::cholesky_optimizedGPU (MatrixOperations.cpp:
 * 124-190), MatrixMatrixOperations::triangularSolve_optimizedGPU
 * (MatrixMatrixOperations.cpp:50-107) on blocks k+1 .. n-1,
 * symmetricMatrixMatrixDiagonal_optimizedGPU (:212-305) on the diagonal
 * blocks k+1 .. n-1 and, at the default gpuOptimizationLevel 3
 * (Configuration.hpp:49), matrixMatrixStep_optimizedGPU3 (:645-779) on the
 * off-diagonal blocks of rows k+2 .. n-1. Every block is written by one
 * kernel per step, so only the order inside one element's computation
 * matters:
 *  - POTRF: right-looking by column; a_ij (i >= j) takes a_ij - a_ik a_jk for
 *    k = 0 .. j-1 in turn, then a_ij / sqrt(a_jj) (i > j) with the diagonal's
 *    own updates done, and a_jj = sqrt(a_jj). Rows >= N are not touched; the
 *    upper triangle of the other rows is zeroed. This is or_potrf's order.
 *  - TRSM, one row x of each block below: d_i = 1.0 / l_ii; v_i starts at the
 *    row's value and takes v_i - l_ik x_k for k = 0 .. i-1 in turn, where
 *    x_k = v_k d_k; the stored value is v_i d_i. The multiplication by the
 *    reciprocal is where this order differs from the CPU kernel (or_trsm,
 *    which divides).
 *  - SYRK: v = a_ij (i >= j), then v - c_ik c_jk for k = 0 .. b-1 in turn
 *    (or_syrk's order).
 *  - GEMM: s = 0, s + b_ik c_jk for k = 0 .. b-1 in turn, then a_ij - s
 *    (or_gemm's order).
 * As for the CG, what the source leaves open is whether the compiler
 * contracts those multiply-subtract / multiply-add statements; fma = 0 keeps
 * every product rounded on its own (this file is built with
 * -ffp-contract=off), fma = 1 contracts exactly the four `v - x y` /
 * `s + x y` forms above with fma().
 * ------------------------------------------------------------------------- */
static inline double msub(double v, double x, double y, int fma_mode) {
    return fma_mode ? fma(-x, y, v) : v - x * y;
}

void or_potrf_gpu_order(double* A, i64 N, int b, int blockID, int blockRow, int fma_mode) {
    double* B = A + (i64)blockID * b * b;
    for (int k = 0; k < b; ++k) {
        const double sqrtDiag = sqrt(B[(i64)k * b + k]);
        for (int i = k + 1; i < b; ++i)
            if ((i64)blockRow * b + i < N) B[(i64)i * b + k] = B[(i64)i * b + k] / sqrtDiag;
        for (int i = k + 1; i < b; ++i) {
            if ((i64)blockRow * b + i >= N) continue;
            for (int j = k + 1; j <= i; ++j)
                B[(i64)i * b + j] = msub(B[(i64)i * b + j], B[(i64)i * b + k], B[(i64)j * b + k], fma_mode);
        }
        B[(i64)k * b + k] = sqrtDiag;
    }
    for (int i = 0; i < b; ++i) {
        if ((i64)blockRow * b + i >= N) continue;
        for (int j = i + 1; j < b; ++j) B[(i64)i * b + j] = 0;
    }
}

void or_trsm_gpu_order(double* A, int b, int blockID, int blockRow, int blockStart, int blockCount, int fma_mode) {
    const double* L = A + (i64)blockID * b * b;
    double* dinv = (double*)malloc(sizeof(double) * b);
    for (int i = 0; i < b; ++i) dinv[i] = 1.0 / L[(i64)i * b + i];
#pragma omp parallel for collapse(2) schedule(static)
    for (int g = 0; g < blockCount; ++g) {
        for (int r = 0; r < b; ++r) {
            double* row = A + ((i64)blockID + (blockStart - blockRow) + g) * b * b + (i64)r * b;
            for (int i = 0; i < b; ++i) {
                double v = row[i];
                for (int k = 0; k < i; ++k) v = msub(v, L[(i64)i * b + k], row[k], fma_mode);
                row[i] = v * dinv[i];
            }
        }
    }
    free(dinv);
}

void or_syrk_gpu_order(double* A, int b, int blockID, int blockRow, int blockStart, int blockCount, int n,
                       int fma_mode) {
#pragma omp parallel for schedule(static)
    for (int g = 0; g < blockCount; ++g) {
        const int diag = blockStart + g;
        double* D = A + or_block_id(diag, diag, n) * b * b;
        const double* C = A + ((i64)blockID + (blockStart - blockRow) + g) * b * b;
        for (int i = 0; i < b; ++i)
            for (int j = 0; j <= i; ++j) {
                double v = D[(i64)i * b + j];
                for (int k = 0; k < b; ++k) v = msub(v, C[(i64)i * b + k], C[(i64)j * b + k], fma_mode);
                D[(i64)i * b + j] = v;
            }
    }
}

void or_gemm_gpu_order(double* A, int b, int blockID, int blockRow, int blockStart, int blockCount, int n,
                       int fma_mode) {
    for (int I = blockStart; I < blockStart + blockCount; ++I) {
        const double* Bm = A + ((i64)blockID + (I - blockRow)) * b * b;
#pragma omp parallel for schedule(static)
        for (int J = blockRow + 1; J < I; ++J) {
            double* C = A + or_block_id(I, J, n) * b * b;
            const double* Cm = A + ((i64)blockID + (J - blockRow)) * b * b;
            for (int i = 0; i < b; ++i)
                for (int j = 0; j < b; ++j) {
                    double s = 0;
                    for (int k = 0; k < b; ++k)
                        s = fma_mode ? fma(Bm[(i64)i * b + k], Cm[(i64)j * b + k], s)
                                     : s + Bm[(i64)i * b + k] * Cm[(i64)j * b + k];
                    C[(i64)i * b + j] = C[(i64)i * b + j] - s;
                }
        }
    }
}

void or_cholesky_gpu_order(double* A, i64 N, int b, int fma_mode) {
    const int n = (int)((N + b - 1) / b);
    for (int k = 0; k < n; ++k) {
        const int blockID = (int)or_block_id(k, k, n);
        or_potrf_gpu_order(A, N, b, blockID, k, fma_mode);
        const int cnt = n - (k + 1);
        if (cnt > 0) {
            or_trsm_gpu_order(A, b, blockID, k, k + 1, cnt, fma_mode);
            or_syrk_gpu_order(A, b, blockID, k, k + 1, cnt, n, fma_mode);
        }
        if (cnt > 1) or_gemm_gpu_order(A, b, blockID, k, k + 2, cnt - 1, n, fma_mode);
    }
}

/* ---------------------------------------------------------------------------
 * The triangular solves after the factor, in the order of the reference's
 * kernels with the contraction choice of the factor above:
 * TriangularSystemSolver::solve (TriangularSystemSolver.cpp:15-105) runs the
 * same two kernels on the GPU queue at share 1.0 as on the CPU queue:
 * triangularSolveBlockVector (MatrixVectorOperations.cpp:282-335; b_k =
 * b_k / a_kk, then b_i - a_ik b_k) and matrixVectorColumnUpdate (:337-377;
 * s = 0, s + a_ik b_k, then b_i - s). At fma = 0 this is or_tri_solve.
 * ------------------------------------------------------------------------- */
static void trsv_block_fma(const double* A, double* x, i64 N, int b, int blockRow, int blockID, int transposed,
                           int fma_mode) {
    const double* L = A + (i64)blockID * b * b;
    double* xb = x + (i64)blockRow * b;
    for (int s = 0; s < b; ++s) {
        const int k = transposed ? b - (s + 1) : s;
        const double bk = xb[k] / L[(i64)k * b + k];
        if ((i64)blockRow * b + k >= N) continue;
        xb[k] = bk;
        for (int i = 0; i < b; ++i) {
            const int cond = transposed ? (i < k) : (i > k);
            if (!cond) continue;
            xb[i] = msub(xb[i], transposed ? L[(i64)k * b + i] : L[(i64)i * b + k], bk, fma_mode);
        }
    }
}

static void col_update_fma(const double* A, double* x, int b, int blockStart, int blockCount, int blockRow,
                           int blockID, int n, int transposed, int fma_mode) {
    const i64 total = (i64)n * (n + 1) / 2;
#pragma omp parallel for schedule(static)
    for (int g = 0; g < blockCount; ++g) {
        i64 id;
        if (!transposed) id = (i64)blockID + (blockStart - blockRow + g);
        else {
            const i64 m = n - g - blockStart;
            id = total - m * (m + 1) / 2 + blockRow - g - blockStart;
        }
        const double* M = A + id * b * b;
        const double* x0 = x + (i64)blockRow * b;
        double* xi = x + (i64)(blockStart + g) * b;
        for (int li = 0; li < b; ++li) {
            double sum = 0.0;
            for (int k = 0; k < b; ++k) {
                const double a = transposed ? M[(i64)k * b + li] : M[(i64)li * b + k];
                sum = fma_mode ? fma(a, x0[k], sum) : sum + a * x0[k];
            }
            xi[li] = xi[li] - sum;
        }
    }
}

void or_tri_solve_gpu_order(const double* L, double* x, i64 N, int b, int fma_mode) {
    const int n = (int)((N + b - 1) / b);
    for (int j = 0; j < n; ++j) {
        const int id = (int)or_block_id(j, j, n);
        trsv_block_fma(L, x, N, b, j, id, 0, fma_mode);
        if (j < n - 1) col_update_fma(L, x, b, j + 1, n - (j + 1), j, id, n, 0, fma_mode);
    }
    for (int j = n - 1; j >= 0; --j) {
        const int id = (int)or_block_id(j, j, n);
        trsv_block_fma(L, x, N, b, j, id, 1, fma_mode);
        if (j > 0) col_update_fma(L, x, b, 0, j, j, id, n, 1, fma_mode);
    }
}

/* ---------------------------------------------------------------------------
 * RBF kernel matrix, reference compatibility mode: features are the 8-lag
 * window of a 1-D series t (7 leading zeros), sigma_f^2 = 1, l = 1,
 * sigma_n^2 = 0.01. Reference: MatrixGenerator::generateSPDMatrix,
 * src/utility/MatrixGenerator.cpp:73-162 (distance accumulated in double in
 * lag order, then vl * exp(-0.5/l^2 * dist), noise on i == j; entries outside
 * N are left as they are, i.e. zero padding).
 * ------------------------------------------------------------------------- */
void or_rbf_window(const double* t, i64 N, int b, double* A) {
    const int n = (int)((N + b - 1) / b);
    const int lags = 8;
    double* tin = (double*)calloc(N + lags - 1, sizeof(double));
    for (i64 i = 0; i < N; ++i) tin[i + lags - 1] = t[i];
#pragma omp parallel for schedule(dynamic)
    for (int bi = 0; bi < n; ++bi) {
        for (int bj = 0; bj <= bi; ++bj) {
            double* B = A + or_block_id(bi, bj, n) * b * b;
            for (int il = 0; il < b; ++il) {
                for (int jl = 0; jl < b; ++jl) {
                    const i64 ig = (i64)b * bi + il, jg = (i64)b * bj + jl;
                    if (ig >= N || jg >= N) continue;
                    double dist = 0.0;
                    for (int k = 0; k < lags; ++k) {
                        const double tmp = tin[ig + k] - tin[jg + k];
                        dist += tmp * tmp;
                    }
                    double c = 1.0 * exp(-0.5 / (1.0 * 1.0) * dist);
                    if (ig == jg) c += 0.01;
                    B[(i64)il * b + jl] = c;
                }
            }
        }
    }
    free(tin);
}

/* General RBF: K_ij = sf2 * exp(-||x_i - x_j||^2 / (2 l^2)) + sn2 [i == j],
 * X row-major N x d. Same per-entry operation order as the compatibility
 * mode above (distance summed in feature order, exp of -0.5/l^2 * dist). */
void or_rbf(const double* X, i64 N, int d, int b, double ell, double sf2, double sn2, double* A) {
    const int n = (int)((N + b - 1) / b);
    const double scale = -0.5 / (ell * ell);
#pragma omp parallel for schedule(dynamic)
    for (int bi = 0; bi < n; ++bi) {
        for (int bj = 0; bj <= bi; ++bj) {
            double* B = A + or_block_id(bi, bj, n) * b * b;
            for (int il = 0; il < b; ++il) {
                for (int jl = 0; jl < b; ++jl) {
                    const i64 ig = (i64)b * bi + il, jg = (i64)b * bj + jl;
                    if (ig >= N || jg >= N) continue;
                    double dist = 0.0;
                    for (int k = 0; k < d; ++k) {
                        const double tmp = X[ig * d + k] - X[jg * d + k];
                        dist += tmp * tmp;
                    }
                    double c = sf2 * exp(scale * dist);
                    if (ig == jg) c += sn2;
                    B[(i64)il * b + jl] = c;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * GP prediction. Reference: MatrixGenerator::generateTestKernelMatrix,
 * src/utility/MatrixGenerator.cpp:163-213 (features: the 8-lag windows that
 * readInputVector, :257-272, builds with 7 leading zeros; per entry the
 * distance is summed in lag order as tmp = train - test, then
 * vl * exp(-0.5 / (l * l) * dist), stored transposed at K*[j N + i]), and
 * MatrixVectorOperations::matrixVectorGP, src/kernel/MatrixVectorOperations.cpp:
 * 379-395 (one work item per row, k ascending). or_rbf_cross takes general
 * row-major features and (ell, sf2) with the same per-entry operation order;
 * the reference fixes vl = l = 1.
 * ------------------------------------------------------------------------- */
void or_window_features(const double* t, i64 N, double* X) {
    const int lags = 8;
    for (i64 i = 0; i < N; ++i)
        for (int k = 0; k < lags; ++k) {
            const i64 src = i + k - (lags - 1);
            X[i * lags + k] = src >= 0 ? t[src] : 0.0;
        }
}

void or_rbf_cross(const double* Xtr, i64 N, const double* Xte, i64 M, int d, double ell, double sf2, double* K) {
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < M; ++j) {
        for (i64 i = 0; i < N; ++i) {
            double dist = 0.0;
            for (int k = 0; k < d; ++k) {
                const double tmp = Xtr[i * d + k] - Xte[j * d + k];
                dist += tmp * tmp;
            }
            K[j * N + i] = sf2 * exp(-0.5 / (ell * ell) * dist);
        }
    }
}

void or_gemv_gp(const double* A, const double* b, double* result, i64 n, i64 m) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m; ++i) {
        double sum = 0.0;
        for (i64 k = 0; k < n; ++k) sum += A[i * n + k] * b[k];
        result[i] = sum;
    }
}

/* ---------------------------------------------------------------------------
 * Synthetic GP inputs (the measurement workload's definition, no reference
 * counterpart): SplitMix64 stream from `seed`; uniforms u = (z >> 11) 2^-53;
 * X = 2u - 1 drawn first (row major N x d), then N normals by Box-Muller
 * (two uniforms each: sqrt(-2 ln(1-u1)) cos(2 pi u2)); y_i = sin(sum_k x_ik)
 * + 0.1 xi_i.
 * ------------------------------------------------------------------------- */
static uint64_t sm64_next(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double sm64_unif(uint64_t* s) { return (double)(sm64_next(s) >> 11) * (1.0 / 9007199254740992.0); }

void or_synth_gp(uint64_t seed, i64 N, int d, double* X, double* y) {
    uint64_t s = seed;
    for (i64 i = 0; i < N * d; ++i) X[i] = 2.0 * sm64_unif(&s) - 1.0;
    const double two_pi = 6.283185307179586476925286766559;
    for (i64 i = 0; i < N; ++i) {
        const double u1 = sm64_unif(&s), u2 = sm64_unif(&s);
        const double xi = sqrt(-2.0 * log(1.0 - u1)) * cos(two_pi * u2);
        double sum = 0.0;
        for (int k = 0; k < d; ++k) sum += X[i * d + k];
        y[i] = sin(sum) + 0.1 * xi;
    }
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
