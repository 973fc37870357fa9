// Reference-order blocked Cholesky: the factor of the reference's share-1.0
// GPU path rounding for rounding, for parity checks against the oracle's
// restatement of it (no counterpart kernel of its own; the factorisation is
// Cholesky::solve_heterogeneous at share 1.0, Cholesky.cpp:413-493).
//
// Per step k the reference runs four kernels, each writing disjoint blocks,
// so only the order inside one element's computation fixes the result:
//  - POTRF (MatrixOperations::cholesky_optimizedGPU, MatrixOperations.cpp:
//    124-190): a_ij (i >= j) takes a_ij - a_ik a_jk for k = 0 .. j-1 in turn,
//    then a_ij / sqrt(a_jj) (i > j), a_jj = sqrt(a_jj); rows >= N untouched,
//    the upper triangle of the others zeroed.
//  - TRSM (triangularSolve_optimizedGPU, MatrixMatrixOperations.cpp:50-107):
//    per row, v_i takes v_i - l_ik x_k for k = 0 .. i-1 in turn and
//    x_i = v_i * (1.0 / l_ii).
//  - SYRK (symmetricMatrixMatrixDiagonal_optimizedGPU, :212-305): v = a_ij,
//    then v - c_ik c_jk for k = 0 .. b-1.
//  - GEMM (matrixMatrixStep_optimizedGPU3, :645-779): s = 0, s + b_ik c_jk for
//    k = 0 .. b-1, then a_ij - s.
// The kernels below keep exactly those sequences, in their own layout: the
// POTRF and the TRSM run left-looking (one thread per row, the terms of an
// element applied in the same increasing k), SYRK and GEMM are 64 x 64 LDS
// tiles with a 4 x 4 register tile per thread and k in increasing order.
// FMA selects whether the `v - x y` / `s + x y` statements are contracted
// (one rounding) or not (two): the source leaves that to the compiler.
// The solves that follow (TriangularSystemSolver::solve, TriangularSystem
// Solver.cpp:15-105, the same two kernels on either queue) keep the order of
// triangularSolveBlockVector (MatrixVectorOperations.cpp:282-335: b_k =
// b_k / a_kk, then b_i - a_ik b_k) and matrixVectorColumnUpdate (:337-377:
// s = 0, s + a_ik b_k, then b_i - s).
// Speed is secondary here; the fast paths are hs_cholesky_f64 and
// hs_tri_solve_f64.
#include <algorithm>

#include "hs_internal.h"

namespace {

template <bool FMA>
__device__ __forceinline__ double msub(double v, double x, double y) {
    if constexpr (FMA) {
        return __builtin_fma(-x, y, v);
    } else {
#pragma clang fp contract(off)
        const double p = x * y;
        return v - p;
    }
}

template <bool FMA>
__device__ __forceinline__ double madd(double s, double x, double y) {
    if constexpr (FMA) {
        return __builtin_fma(x, y, s);
    } else {
#pragma clang fp contract(off)
        const double p = x * y;
        return s + p;
    }
}

__device__ __forceinline__ double sub_nc(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}

// One work-group on diagonal block blockID, thread i = row i (b <= 1024),
// column by column: the element's terms, then the division by the root of
// the column's diagonal, which thread j publishes in LDS.
template <bool FMA>
__global__ __launch_bounds__(1024) void potrf_ref_kernel(double* __restrict__ A, i64 N, int b, i64 blockID,
                                                         int blockRow, int* __restrict__ info) {
    __shared__ double piv;
    double* B = A + blockID * (i64)b * b;
    const int i = threadIdx.x;
    const i64 rowBase = (i64)blockRow * b;
    const bool valid = i < b && rowBase + i < N;
    double* Ri = B + (i64)(i < b ? i : 0) * b;
    for (int j = 0; j < b; ++j) {
        const double* Rj = B + (i64)j * b;
        double v = 0.0;
        if (valid && i >= j) {
            v = Ri[j];
            for (int k = 0; k < j; ++k) v = msub<FMA>(v, Ri[k], Rj[k]);
        }
        if (valid && i == j) {
            piv = __builtin_sqrt(v);
            if (!(v > 0.0)) atomicCAS(info, 0, (int)(rowBase + j + 1));
        }
        __syncthreads();
        if (valid && i > j) Ri[j] = v / piv;
        if (valid && i == j) Ri[j] = piv;
        __syncthreads();
    }
    if (valid)
        for (int j = i + 1; j < b; ++j) Ri[j] = 0.0;
}

// One thread per row of the blockCount panel blocks: forward substitution
// x L^T = row in place, terms in increasing k; L's entries are the same for
// every row of a wave (broadcast loads), the reciprocals sit in LDS.
template <bool FMA>
__global__ __launch_bounds__(256) void trsm_ref_kernel(double* __restrict__ A, int b, i64 blockID, int blockRow,
                                                       int blockStart, int blockCount) {
    extern __shared__ double dinv[];
    const double* L = A + blockID * (i64)b * b;
    for (int t = threadIdx.x; t < b; t += blockDim.x) dinv[t] = 1.0 / L[(i64)t * b + t];
    __syncthreads();
    const i64 r = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (i64)blockCount * b) return;
    const i64 g = r / b, rr = r % b;
    double* row = A + (blockID + (blockStart - blockRow) + g) * (i64)b * b + rr * b;
    for (int i = 0; i < b; ++i) {
        const double* Li = L + (i64)i * b;
        double v = row[i];
        for (int k = 0; k < i; ++k) v = msub<FMA>(v, Li[k], row[k]);
        row[i] = v * dinv[i];
    }
}

constexpr int RT = 64;  // output tile edge
constexpr int RK = 16;  // k chunk through LDS

// SYRK (GEMM = false): D_I <- D_I - P P^T on the lower triangle of diagonal
// block I = blockStart + blockIdx.y, P = block (I, blockRow).
// GEMM (GEMM = true): A_IJ <- A_IJ - s, s = sum_k L_Ik L_Jk accumulated from
// zero, for block pair blockIdx.y -> (I, J), blockRow < J < I.
// 256 threads as 16 x 16, each owning rows ty + 16 ii and columns tx + 16 jj
// of the 64 x 64 tile blockIdx.x.
template <bool GEMM, bool FMA>
__global__ __launch_bounds__(256) void update_ref_kernel(double* __restrict__ A, int b, int n, i64 blockID,
                                                         int blockRow, int blockStart, int blockCount) {
    __shared__ double Xs[RT][RK + 1];
    __shared__ double Ys[RT][RK + 1];
    const int tpe = (b + RT - 1) / RT;
    const int tm = blockIdx.x / tpe, tn = blockIdx.x % tpe;
    int I, J;
    if constexpr (GEMM) {
        // pair index over I in [blockStart, blockStart + blockCount) and
        // J in [blockRow + 1, blockStart + blockCount - 1); J >= I idle
        const int span = blockStart + blockCount - 1 - (blockRow + 1);
        I = blockStart + (int)(blockIdx.y / span);
        J = blockRow + 1 + (int)(blockIdx.y % span);
        if (J >= I) return;
    } else {
        I = blockStart + (int)blockIdx.y;
        J = I;
        if (tn > tm) return;  // tile wholly above the diagonal
    }
    const i64 bb = (i64)b * b;
    double* C = A + hs_block_id(I, J, n) * bb;
    const double* X = A + (blockID + (I - blockRow)) * bb;  // block (I, k)
    const double* Y = A + (blockID + (J - blockRow)) * bb;  // block (J, k)
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int r0 = tm * RT, c0 = tn * RT;
    double acc[4][4];
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int r = r0 + ty + 16 * ii, c = c0 + tx + 16 * jj;
            acc[ii][jj] = (!GEMM && r < b && c < b && r >= c) ? C[(i64)r * b + c] : 0.0;
        }
    for (int k0 = 0; k0 < b; k0 += RK) {
        for (int e = threadIdx.x; e < RT * RK; e += 256) {
            const int rr = e / RK, kk = e % RK;
            const int k = k0 + kk;
            Xs[rr][kk] = (r0 + rr < b && k < b) ? X[(i64)(r0 + rr) * b + k] : 0.0;
            Ys[rr][kk] = (c0 + rr < b && k < b) ? Y[(i64)(c0 + rr) * b + k] : 0.0;
        }
        __syncthreads();
        const int kn = std::min(RK, b - k0);
        for (int kk = 0; kk < kn; ++kk) {
#pragma unroll
            for (int ii = 0; ii < 4; ++ii) {
                const double x = Xs[ty + 16 * ii][kk];
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const double y = Ys[tx + 16 * jj][kk];
                    acc[ii][jj] = GEMM ? madd<FMA>(acc[ii][jj], x, y) : msub<FMA>(acc[ii][jj], x, y);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            const int r = r0 + ty + 16 * ii, c = c0 + tx + 16 * jj;
            if (r >= b || c >= b) continue;
            if (GEMM) {
                double* p = C + (i64)r * b + c;
                *p = sub_nc(*p, acc[ii][jj]);
            } else if (r >= c) {
                C[(i64)r * b + c] = acc[ii][jj];
            }
        }
}

// One work-group on diagonal block blockID, one thread per entry of x's block
// blockRow, kept in LDS: per step the pivot entry is divided by the diagonal
// and subtracted from the entries below it (above it when transposed).
// Entries >= N are neither written nor used.
template <bool FMA>
__global__ __launch_bounds__(1024) void trsv_ref_kernel(const double* __restrict__ A, double* __restrict__ x,
                                                        i64 N, int b, int blockRow, i64 blockID, int transposed) {
    extern __shared__ double xs[];
    const double* L = A + blockID * (i64)b * b;
    double* xb = x + (i64)blockRow * b;
    const int i = threadIdx.x;
    if (i < b) xs[i] = xb[i];
    __syncthreads();
    for (int s = 0; s < b; ++s) {
        const int k = transposed ? b - (s + 1) : s;
        const double bk = xs[k] / L[(i64)k * b + k];
        __syncthreads();
        if ((i64)blockRow * b + k < N && i < b) {
            if (i == k) xs[k] = bk;
            else if (transposed ? i < k : i > k)
                xs[i] = msub<FMA>(xs[i], transposed ? L[(i64)k * b + i] : L[(i64)i * b + k], bk);
        }
        __syncthreads();
    }
    if (i < b) xb[i] = xs[i];
}

// x_I -= L_{I,blockRow} x_blockRow (transposed: L_{blockRow,I}^T x_blockRow),
// one work-group per block I, one thread per entry: the sum from zero over k
// in order, then the subtraction.
template <bool FMA>
__global__ __launch_bounds__(1024) void colupd_ref_kernel(const double* __restrict__ A, double* __restrict__ x,
                                                          int b, int blockStart, int blockRow, i64 blockID, int n,
                                                          int transposed) {
    const int g = blockIdx.x, li = threadIdx.x;
    if (li >= b) return;
    i64 id;
    if (!transposed) {
        id = blockID + (blockStart - blockRow + g);
    } else {
        const i64 total = (i64)n * (n + 1) / 2, m = n - g - blockStart;
        id = total - m * (m + 1) / 2 + blockRow - g - blockStart;
    }
    const double* M = A + id * (i64)b * b;
    const double* x0 = x + (i64)blockRow * b;
    double sum = 0.0;
    for (int k = 0; k < b; ++k) sum = madd<FMA>(sum, transposed ? M[(i64)k * b + li] : M[(i64)li * b + k], x0[k]);
    double* xi = x + (i64)(blockStart + g) * b + li;
    *xi = sub_nc(*xi, sum);
}

int check_common(hs_handle_t h, const double* A, int b, int fma) {
    if (!h) return -1;
    if (!A) return -2;
    if (b <= 0) return -3;
    if (fma != 0 && fma != 1) {
        hs_set_error("reference-order factor: fma must be 0 or 1");
        return HS_ERR_UNSUPPORTED;
    }
    return HS_OK;
}

int potrf_ref(hs_handle_t h, double* A, i64 N, int b, i64 blockID, int blockRow, int fma) {
    if (b > 1024) {
        hs_set_error("reference-order POTRF: matrix_bsz must be <= 1024");
        return HS_ERR_UNSUPPORTED;
    }
    const int threads = (b + 63) / 64 * 64;
    if (fma) potrf_ref_kernel<true><<<1, threads, 0, h->stream>>>(A, N, b, blockID, blockRow, h->info_dev);
    else potrf_ref_kernel<false><<<1, threads, 0, h->stream>>>(A, N, b, blockID, blockRow, h->info_dev);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

int trsm_ref(hs_handle_t h, double* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount, int fma) {
    if (blockCount <= 0) return HS_OK;
    const i64 rows = (i64)blockCount * b;
    const unsigned grid = (unsigned)((rows + 255) / 256);
    const size_t lds = sizeof(double) * (size_t)b;
    if (lds > 64 * 1024) {
        hs_set_error("reference-order TRSM: matrix_bsz too large");
        return HS_ERR_UNSUPPORTED;
    }
    if (fma) trsm_ref_kernel<true><<<grid, 256, lds, h->stream>>>(A, b, blockID, blockRow, blockStart, blockCount);
    else trsm_ref_kernel<false><<<grid, 256, lds, h->stream>>>(A, b, blockID, blockRow, blockStart, blockCount);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

int syrk_ref(hs_handle_t h, double* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount, int n,
             int fma) {
    if (blockCount <= 0) return HS_OK;
    const int tpe = (b + RT - 1) / RT;
    const dim3 grid((unsigned)(tpe * tpe), (unsigned)blockCount);
    if (fma) update_ref_kernel<false, true><<<grid, 256, 0, h->stream>>>(A, b, n, blockID, blockRow, blockStart, blockCount);
    else update_ref_kernel<false, false><<<grid, 256, 0, h->stream>>>(A, b, n, blockID, blockRow, blockStart, blockCount);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

int gemm_ref(hs_handle_t h, double* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount, int n,
             int fma) {
    const int span = blockStart + blockCount - 1 - (blockRow + 1);
    if (blockCount <= 0 || span <= 0) return HS_OK;
    const i64 pairs = (i64)blockCount * span;
    if (pairs > 65535) {
        hs_set_error("reference-order GEMM: too many block pairs for one launch");
        return HS_ERR_UNSUPPORTED;
    }
    const int tpe = (b + RT - 1) / RT;
    const dim3 grid((unsigned)(tpe * tpe), (unsigned)pairs);
    if (fma) update_ref_kernel<true, true><<<grid, 256, 0, h->stream>>>(A, b, n, blockID, blockRow, blockStart, blockCount);
    else update_ref_kernel<true, false><<<grid, 256, 0, h->stream>>>(A, b, n, blockID, blockRow, blockStart, blockCount);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// the block ranges of the reference's kernels must stay inside the packed
// matrix of n block columns
bool ranges_ok(int n, int blockID, int blockRow, int blockStart, int blockCount) {
    return n > 0 && blockRow >= 0 && blockRow < n && blockID == (int)hs_block_id(blockRow, blockRow, n) &&
           blockStart > blockRow && blockCount >= 0 && blockStart + blockCount <= n;
}

}  // namespace

extern "C" int hs_potrf_ref_order_f64(hs_handle_t h, double* A, int64_t N, int b, int blockID, int blockRow,
                                      int fma) {
    int rc = check_common(h, A, b, fma);
    if (rc) return rc;
    if (N <= 0 || blockID < 0 || blockRow < 0) return -4;
    return potrf_ref(h, A, N, b, blockID, blockRow, fma);
}

extern "C" int hs_trsm_ref_order_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                                     int blockCount, int fma) {
    int rc = check_common(h, A, b, fma);
    if (rc) return rc;
    if (blockID < 0 || blockRow < 0 || blockStart <= blockRow || blockCount < 0) return -4;
    return trsm_ref(h, A, b, blockID, blockRow, blockStart, blockCount, fma);
}

extern "C" int hs_syrk_ref_order_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                                     int blockCount, int blockCountXY, int fma) {
    int rc = check_common(h, A, b, fma);
    if (rc) return rc;
    if (!ranges_ok(blockCountXY, blockID, blockRow, blockStart, blockCount)) return -4;
    return syrk_ref(h, A, b, blockID, blockRow, blockStart, blockCount, blockCountXY, fma);
}

extern "C" int hs_gemm_ref_order_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                                     int blockCount, int blockCountXY, int fma) {
    int rc = check_common(h, A, b, fma);
    if (rc) return rc;
    if (blockCount > 0 && (!ranges_ok(blockCountXY, blockID, blockRow, blockStart, blockCount) ||
                           blockStart < blockRow + 2))
        return -4;
    return gemm_ref(h, A, b, blockID, blockRow, blockStart, blockCount, blockCountXY, fma);
}

extern "C" int hs_cholesky_ref_order_f64(hs_handle_t h, double* A, int64_t N, int b, int fma) {
    int rc = check_common(h, A, b, fma);
    if (rc) return rc;
    if (N <= 0) return -4;
    const int n = (int)((N + b - 1) / b);
    if ((i64)(n - 2) * (n - 2) > 65535 && n > 2) {
        // the widest GEMM launch (step 0) has (n - 2)^2 block pairs
        hs_set_error("hs_cholesky_ref_order_f64: at most 257 block columns");
        return HS_ERR_UNSUPPORTED;
    }
    if ((rc = hs_reset_info(h))) return rc;
    for (int k = 0; k < n; ++k) {
        const i64 id = hs_block_id(k, k, n);
        if ((rc = potrf_ref(h, A, N, b, id, k, fma))) return rc;
        const int cnt = n - (k + 1);
        if (cnt > 0) {
            if ((rc = trsm_ref(h, A, b, id, k, k + 1, cnt, fma))) return rc;
            if ((rc = syrk_ref(h, A, b, id, k, k + 1, cnt, n, fma))) return rc;
        }
        if (cnt > 1 && (rc = gemm_ref(h, A, b, id, k, k + 2, cnt - 1, n, fma))) return rc;
    }
    int64_t info = 0;
    if ((rc = hs_get_info(h, &info))) return rc;
    return (int)info;
}

extern "C" int hs_tri_solve_ref_order_f64(hs_handle_t h, const double* L, double* x, int64_t N, int b, int fma) {
    int rc = check_common(h, L, b, fma);
    if (rc) return rc;
    if (!x) return -3;
    if (N <= 0) return -4;
    if (b > 1024) {
        hs_set_error("hs_tri_solve_ref_order_f64: matrix_bsz must be <= 1024");
        return HS_ERR_UNSUPPORTED;
    }
    const int n = (int)((N + b - 1) / b);
    const int threads = (b + 63) / 64 * 64;
    const size_t lds = sizeof(double) * (size_t)b;
    hipStream_t s = h->stream;
    auto trsv = [&](int j, int tr) {
        const i64 id = hs_block_id(j, j, n);
        if (fma) trsv_ref_kernel<true><<<1, threads, lds, s>>>(L, x, N, b, j, id, tr);
        else trsv_ref_kernel<false><<<1, threads, lds, s>>>(L, x, N, b, j, id, tr);
    };
    auto colupd = [&](int start, int count, int j, int tr) {
        const i64 id = hs_block_id(j, j, n);
        if (fma) colupd_ref_kernel<true><<<count, threads, 0, s>>>(L, x, b, start, j, id, n, tr);
        else colupd_ref_kernel<false><<<count, threads, 0, s>>>(L, x, b, start, j, id, n, tr);
    };
    for (int j = 0; j < n; ++j) {
        trsv(j, 0);
        if (j < n - 1) colupd(j + 1, n - (j + 1), j, 0);
    }
    for (int j = n - 1; j >= 0; --j) {
        trsv(j, 1);
        if (j > 0) colupd(0, j, j, 1);
    }
    HS_LAUNCH_CHECK();
    return HS_OK;
}
