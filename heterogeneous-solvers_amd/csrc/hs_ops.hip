// Reference-shaped kernel ops (one C-ABI entry per reference L3 kernel).
// These kernels take any block size b (the reference's tests use 4, 6, 9, 20,
// 24); the solver drivers use the MFMA / streaming kernels in hs_chol.hip and
// hs_cg.hip where b allows it, and these where it does not.
#include <algorithm>

#include "hs_internal.h"

// ---------------------------------------------------------------------------
// POTRF of one diagonal block, single work-group, left-looking over panels of
// PW columns held in LDS. Rows >= N (padding) are neither read-modified nor
// written; the upper triangle of valid rows is zeroed, as
// MatrixOperations::cholesky_optimizedGPU does (MatrixOperations.cpp:177-186).
// The first non-positive pivot of a valid column is recorded in *info
// (1-based global row); the reference silently takes sqrt of it
// (MatrixOperations.cpp:149).
// LDS: P[b][PW+1] panel + T[b][KT+1] staged k-chunk of finished columns.
// ---------------------------------------------------------------------------
constexpr int POTRF_PW = 16;
constexpr int POTRF_KT = 16;
constexpr int POTRF_THREADS = 256;

__global__ __launch_bounds__(POTRF_THREADS) void potrf_block_kernel(double* __restrict__ A, i64 N, int b, i64 blockID,
                                                                     int blockRow, int* __restrict__ info) {
    extern __shared__ double smem[];
    double* P = smem;                               // [b][PW+1]
    double* T = smem + (size_t)b * (POTRF_PW + 1);  // [b][KT+1]
    double* B = A + blockID * (i64)b * b;
    const int tid = threadIdx.x;
    const i64 rowBase = (i64)blockRow * b;
    const int nvalid = (int)std::min<i64>((i64)b, std::max<i64>((i64)0, N - rowBase));

    for (int p = 0; p < b; p += POTRF_PW) {
        const int pw = std::min(POTRF_PW, b - p);
        const int rows = b - p;
        for (int e = tid; e < rows * POTRF_PW; e += POTRF_THREADS) {
            const int r = e / POTRF_PW, c = e % POTRF_PW;
            double v = 0.0;
            if (c < pw && r >= c) v = B[(i64)(p + r) * b + p + c];
            P[r * (POTRF_PW + 1) + c] = v;
        }
        __syncthreads();
        // left-looking update with the finished columns [0, p)
        for (int k0 = 0; k0 < p; k0 += POTRF_KT) {
            const int kt = std::min(POTRF_KT, p - k0);
            for (int e = tid; e < rows * POTRF_KT; e += POTRF_THREADS) {
                const int r = e / POTRF_KT, kk = e % POTRF_KT;
                T[r * (POTRF_KT + 1) + kk] = (kk < kt) ? B[(i64)(p + r) * b + k0 + kk] : 0.0;
            }
            __syncthreads();
            for (int e = tid; e < rows * POTRF_PW; e += POTRF_THREADS) {
                const int r = e / POTRF_PW, c = e % POTRF_PW;
                if (c < pw && r >= c && p + r < nvalid) {
                    double acc = 0.0;
                    for (int kk = 0; kk < kt; ++kk) acc += T[r * (POTRF_KT + 1) + kk] * T[c * (POTRF_KT + 1) + kk];
                    P[r * (POTRF_PW + 1) + c] -= acc;
                }
            }
            __syncthreads();
        }
        // unblocked right-looking factorisation of the tall panel
        for (int c = 0; c < pw; ++c) {
            const double d = P[c * (POTRF_PW + 1) + c];
            const double s = sqrt(d);
            const bool colValid = (p + c) < nvalid;
            if (tid == 0 && colValid && !(d > 0.0)) atomicCAS(info, 0, (int)(rowBase + p + c + 1));
            for (int r = c + 1 + tid; r < rows; r += POTRF_THREADS)
                if (p + r < nvalid) P[r * (POTRF_PW + 1) + c] /= s;
            __syncthreads();
            if (tid == 0) P[c * (POTRF_PW + 1) + c] = s;
            for (int e = tid; e < rows * POTRF_PW; e += POTRF_THREADS) {
                const int r = e / POTRF_PW, cc = e % POTRF_PW;
                if (cc > c && cc < pw && r >= cc && p + r < nvalid)
                    P[r * (POTRF_PW + 1) + cc] -= P[r * (POTRF_PW + 1) + c] * P[cc * (POTRF_PW + 1) + c];
            }
            __syncthreads();
        }
        // write the panel back; zero the upper triangle of valid rows
        for (int e = tid; e < rows * POTRF_PW; e += POTRF_THREADS) {
            const int r = e / POTRF_PW, c = e % POTRF_PW;
            if (c < pw) {
                const int gi = p + r;
                if (gi < nvalid) B[(i64)gi * b + p + c] = (r >= c) ? P[r * (POTRF_PW + 1) + c] : 0.0;
                else if (r == c) B[(i64)gi * b + p + c] = P[r * (POTRF_PW + 1) + c];
            }
        }
        for (int e = tid; e < pw * b; e += POTRF_THREADS) {
            const int r = e / b, j = e % b;
            if (j >= p + pw && p + r < nvalid) B[(i64)(p + r) * b + j] = 0.0;
        }
        __syncthreads();
    }
}

// Panel kernel (hs_panel.hip) for the solver block sizes, else the generic one.
int hs_launch_potrf(hs_handle_t h, double* A, i64 N, int b, i64 blockID, int blockRow) {
    if (hs_potrf_fast_ok(b)) return hs_launch_potrf_fast<double>(h, A, N, b, blockID, blockRow, nullptr);
    const size_t lds = (size_t)b * (POTRF_PW + 1 + POTRF_KT + 1) * sizeof(double);
    if (lds > 160 * 1024) {
        hs_set_error("POTRF: block size too large for one work-group's LDS (b <= 512)");
        return HS_ERR_UNSUPPORTED;
    }
    if (lds > 64 * 1024) {
        HS_HIP(hipFuncSetAttribute((const void*)potrf_block_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    }
    potrf_block_kernel<<<1, POTRF_THREADS, lds, h->stream>>>(A, N, b, blockID, blockRow, h->info_dev);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// ---------------------------------------------------------------------------
// Panel TRSM, B <- B L^{-T}: a work-group owns TR rows of one panel block;
// thread t owns column t of those rows in registers. Step c: the owner of
// column c divides by l_cc and publishes the solved column in LDS, every later
// column subtracts x_c * l_tc. Per element this is the operation order of
// triangularSolve_optimizedCPU (MatrixMatrixOperations.cpp:128-139).
// ---------------------------------------------------------------------------
constexpr int TRSM_TR = 32;

__global__ __launch_bounds__(1024) void trsm_panel_kernel(double* __restrict__ A, int b, i64 blockID, int blockRow,
                                                           int blockStart, int strips) {
    __shared__ double xs[2][TRSM_TR];
    const int g = blockIdx.x / strips;
    const int s = blockIdx.x % strips;
    const int r0 = s * TRSM_TR;
    const int t = threadIdx.x;
    const double* L = A + blockID * (i64)b * b;
    double* Bm = A + (blockID + (blockStart - blockRow) + g) * (i64)b * b;
    double x[TRSM_TR];
#pragma unroll
    for (int i = 0; i < TRSM_TR; ++i) x[i] = (t < b && r0 + i < b) ? Bm[(i64)(r0 + i) * b + t] : 0.0;
    for (int c = 0; c < b; ++c) {
        const int buf = c & 1;
        if (t == c) {
            const double lcc = L[(i64)c * b + c];
#pragma unroll
            for (int i = 0; i < TRSM_TR; ++i) {
                x[i] = x[i] / lcc;
                xs[buf][i] = x[i];
            }
        }
        __syncthreads();
        if (t > c && t < b) {
            const double ltc = L[(i64)t * b + c];
#pragma unroll
            for (int i = 0; i < TRSM_TR; ++i) x[i] = x[i] - ltc * xs[buf][i];
        }
    }
    if (t < b) {
#pragma unroll
        for (int i = 0; i < TRSM_TR; ++i)
            if (r0 + i < b) Bm[(i64)(r0 + i) * b + t] = x[i];
    }
}

int hs_launch_trsm(hs_handle_t h, double* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount) {
    if (blockCount <= 0) return HS_OK;
    if (b > 1024) {
        hs_set_error("TRSM: block size above 1024 is not supported");
        return HS_ERR_UNSUPPORTED;
    }
    if (hs_trsm_fast_ok(b)) return hs_launch_trsm_fast<double>(h, A, b, blockID, blockRow, blockStart, blockCount, nullptr);
    const int strips = (b + TRSM_TR - 1) / TRSM_TR;
    const int threads = ((b + 63) / 64) * 64;
    trsm_panel_kernel<<<dim3((unsigned)(blockCount * strips)), dim3(threads), 0, h->stream>>>(A, b, blockID, blockRow,
                                                                                             blockStart, strips);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// ---------------------------------------------------------------------------
// Element-wise SYRK / GEMM for block sizes the MFMA kernel does not take.
// One thread per output element: acc = sum_k L_Ik[i][k] L_Jk[j][k], then
// A_IJ[i][j] -= acc (the order of matrixMatrixStep_optimizedGPU3, :757-773).
// mode 1: diagonal blocks (I == J, i >= j only); 2: I > J; 3: both.
// The trailing triangle rows are [rowStart, rowStart+rowCount), columns
// J in [k+1, I] (mode 1: J = I; mode 2: J < I).
// ---------------------------------------------------------------------------
__global__ void trailing_elem_kernel(double* __restrict__ A, int b, int k, int rowStart, int n, int mode, i64 pairs,
                                     int pairOffset) {
    const i64 bb = (i64)b * b;
    const i64 gid = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    const i64 q = gid / bb;
    if (q >= pairs) return;
    const int e = (int)(gid % bb);
    const int i = e / b, j = e % b;
    int I, J;
    if (mode == 1) {
        I = rowStart + (int)q;
        J = I;
    } else {
        // rows of the trailing triangle starting at k+1; pairs counted from pairOffset
        int r, c;
        hs_tri_decode(q + pairOffset, &r, &c);
        I = k + 1 + r;
        J = k + 1 + c;
        if (mode == 2) {
            // strictly lower pairs: decode over (I-k-2, J-k-1) lower-with-diagonal triangle
            I = k + 2 + r;
            J = k + 1 + c;
        }
    }
    if (I == J && i < j) return;
    const double* Li = A + hs_block_id(I, k, n) * bb;
    const double* Lj = A + hs_block_id(J, k, n) * bb;
    double acc = 0.0;
    for (int kk = 0; kk < b; ++kk) acc += Li[(i64)i * b + kk] * Lj[(i64)j * b + kk];
    A[hs_block_id(I, J, n) * bb + (i64)i * b + j] -= acc;
}

int hs_launch_trailing_generic(hs_handle_t h, double* A, int b, int k, int rowStart, int rowCount, int n, int mode) {
    if (rowCount <= 0) return HS_OK;
    const i64 bb = (i64)b * b;
    if (mode & 1) {
        const i64 pairs = rowCount;
        const i64 threads = pairs * bb;
        trailing_elem_kernel<<<dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, h->stream>>>(A, b, k, rowStart, n,
                                                                                                   1, pairs, 0);
        HS_LAUNCH_CHECK();
    }
    if (mode & 2) {
        // strictly-lower pairs with I in [rowStart, rowStart+rowCount): I' = I-k-2 >= 0
        const int firstRow = std::max(rowStart, k + 2);
        const int lastRow = rowStart + rowCount;  // exclusive
        if (lastRow > firstRow) {
            const i64 r0 = firstRow - (k + 2), r1 = lastRow - (k + 2);
            const i64 off = r0 * (r0 + 1) / 2;
            const i64 pairs = r1 * (r1 + 1) / 2 - off;
            const i64 threads = pairs * bb;
            trailing_elem_kernel<<<dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, h->stream>>>(
                A, b, k, rowStart, n, 2, pairs, (int)off);
            HS_LAUNCH_CHECK();
        }
    }
    return HS_OK;
}

// ---------------------------------------------------------------------------
// Sub-range symmetric mat-vec, one thread per row, the arithmetic order of
// matrixVectorBlock (MatrixVectorOperations.cpp:8-93).
// ---------------------------------------------------------------------------
__global__ void symv_rows_kernel(const double* __restrict__ A, const double* __restrict__ x, double* __restrict__ y,
                                 int b, int bsi, int bsj, int bci, int bcj, int n, int reset) {
    const i64 li = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= (i64)bci * b) return;
    const i64 i = li + (i64)bsi * b;
    const int iIn = (int)(i % b);
    const int bi = (int)(i / b);
    const i64 bb = (i64)b * b;
    double acc = 0.0;
    if (!reset) acc += y[i];
    int bj = bsj;
    const int last = bsj + bcj - 1;
    for (; bj <= min(bi, last); ++bj) {
        const double* row = A + hs_block_id(bi, bj, n) * bb + (i64)iIn * b;
        const double* xv = x + (i64)bj * b;
        for (int j = 0; j < b; ++j) acc += row[j] * xv[j];
    }
    for (; bj < bsj + bcj; ++bj) {
        const double* blk = A + hs_block_id(bj, bi, n) * bb;
        const double* xv = x + (i64)bj * b;
        for (int j = 0; j < b; ++j) acc += blk[(i64)j * b + iIn] * xv[j];
    }
    y[i] = acc;
}

// ---------------------------------------------------------------------------
// TRSV on one diagonal block: one work-group, x_J staged in LDS, b serial
// steps (triangularSolveBlockVector, MatrixVectorOperations.cpp:282-335).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void trsv_block_kernel(const double* __restrict__ A, double* __restrict__ x, i64 N,
                                                           int b, int blockRow, i64 blockID, int transposed) {
    extern __shared__ double xs[];
    const int t = threadIdx.x;
    const double* L = A + blockID * (i64)b * b;
    double* xb = x + (i64)blockRow * b;
    for (int i = t; i < b; i += blockDim.x) xs[i] = xb[i];
    for (int s = 0; s < b; ++s) {
        const int k = transposed ? b - (s + 1) : s;
        __syncthreads();
        const double bk = xs[k] / L[(i64)k * b + k];
        __syncthreads();
        if ((i64)blockRow * b + k < N) {
            if (t == 0) xs[k] = bk;
            for (int i = t; i < b; i += blockDim.x) {
                if (transposed ? (i < k) : (i > k))
                    xs[i] = xs[i] - (transposed ? L[(i64)k * b + i] : L[(i64)i * b + k]) * bk;
            }
        }
    }
    __syncthreads();
    for (int i = t; i < b; i += blockDim.x) xb[i] = xs[i];
}

int hs_launch_trsv_block(hs_handle_t h, const double* A, double* x, i64 N, int b, int blockRow, i64 blockID,
                         int transposed) {
    const int threads = std::min(1024, ((b + 63) / 64) * 64);
    trsv_block_kernel<<<1, threads, (size_t)b * sizeof(double), h->stream>>>(A, x, N, b, blockRow, blockID,
                                                                             transposed);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// ---------------------------------------------------------------------------
// Column update of the triangular solves: one work-group per target block,
// the input block x_blockRow staged in LDS, one thread per row
// (matrixVectorColumnUpdate, MatrixVectorOperations.cpp:337-377).
// ---------------------------------------------------------------------------
__global__ void col_update_kernel(const double* __restrict__ A, double* __restrict__ x, int b, int blockStart,
                                  int blockRow, i64 blockID, int n, int transposed) {
    extern __shared__ double x0[];
    const int g = blockIdx.x;
    const i64 bb = (i64)b * b;
    for (int i = threadIdx.x; i < b; i += blockDim.x) x0[i] = x[(i64)blockRow * b + i];
    __syncthreads();
    i64 id;
    if (!transposed) {
        id = blockID + (blockStart - blockRow + g);
    } else {
        const i64 total = (i64)n * (n + 1) / 2;
        const i64 m = n - g - blockStart;
        id = total - m * (m + 1) / 2 + blockRow - g - blockStart;
    }
    const double* M = A + id * bb;
    double* xi = x + (i64)(blockStart + g) * b;
    for (int li = threadIdx.x; li < b; li += blockDim.x) {
        double sum = 0.0;
        if (!transposed)
            for (int k = 0; k < b; ++k) sum += M[(i64)li * b + k] * x0[k];
        else
            for (int k = 0; k < b; ++k) sum += M[(i64)k * b + li] * x0[k];
        xi[li] -= sum;
    }
}

int hs_launch_col_update(hs_handle_t h, const double* A, double* x, int b, int blockStart, int blockCount, int blockRow,
                         i64 blockID, int n, int transposed) {
    if (blockCount <= 0) return HS_OK;
    const int threads = std::min(1024, ((b + 63) / 64) * 64);
    col_update_kernel<<<dim3((unsigned)blockCount), dim3(threads), (size_t)b * sizeof(double), h->stream>>>(
        A, x, b, blockStart, blockRow, blockID, n, transposed);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// ---------------------------------------------------------------------------
// Vector block ops (VectorOperations.cpp:6-82) and the two-pass dot product.
// ---------------------------------------------------------------------------
__global__ void vec_op_kernel(const double* __restrict__ x, const double* __restrict__ y, double alpha,
                              double* __restrict__ r, i64 begin, i64 end, int op) {
    for (i64 i = begin + (i64)blockIdx.x * blockDim.x + threadIdx.x; i < end; i += (i64)gridDim.x * blockDim.x) {
        switch (op) {
            case 0: r[i] = x[i] * alpha; break;
            case 1: r[i] = x[i] + y[i]; break;
            case 2: r[i] = x[i] - y[i]; break;
            default: r[i] = x[i] + y[i] * alpha; break;
        }
    }
}

static int launch_vec(hs_handle_t h, const double* x, const double* y, double alpha, double* r, int b, int bs, int bc,
                      int op) {
    if (!h) return -1;
    if (b <= 0) return -5;
    if (bc <= 0) return HS_OK;
    const i64 begin = (i64)bs * b, end = (i64)(bs + bc) * b;
    i64 blocks = (end - begin + 255) / 256;
    blocks = std::min<i64>(blocks, 4096);
    vec_op_kernel<<<dim3((unsigned)blocks), dim3(256), 0, h->stream>>>(x, y, alpha, r, begin, end, op);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// Streaming-read probe: work-group g sums x[g*RS_CHUNK .. +RS_CHUNK) with
// 16-byte non-temporal loads, eight per lane in flight, and writes its sum
// to part[g] (the sums keep the loads live; the figure of interest is the
// read rate). vec = 0 (x not 16-byte aligned) or a partial last chunk reads
// one element at a time.
constexpr int RS_CHUNK = 16384;  // doubles (128 KB) per work-group
typedef double rs_d2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void read_sum_kernel(const double* __restrict__ x, i64 count,
                                                       double* __restrict__ part, int vec) {
    __shared__ double red[4];
    const i64 base = (i64)blockIdx.x * RS_CHUNK;
    const i64 end = min(count, base + RS_CHUNK);
    double s = 0.0;
    if (vec && end - base == RS_CHUNK) {
        const rs_d2* p = reinterpret_cast<const rs_d2*>(x + base);
#pragma unroll
        for (int r = 0; r < RS_CHUNK / 2 / 256 / 8; ++r) {
            rs_d2 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(p + (r * 8 + u) * 256 + threadIdx.x);
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u].x + v[u].y;
        }
    } else {
        for (i64 i = base + threadIdx.x; i < end; i += 256) s += x[i];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// First pass: G work-groups, each sums a contiguous chunk in a fixed order
// (per-thread strided partial sums, then a fixed LDS tree).
constexpr int DOT_THREADS = 256;

__global__ __launch_bounds__(DOT_THREADS) void dot_partial_kernel(const double* __restrict__ x,
                                                                   const double* __restrict__ y, i64 begin, i64 len,
                                                                   i64 chunk, double* __restrict__ part) {
    __shared__ double red[DOT_THREADS];
    const i64 s = begin + (i64)blockIdx.x * chunk;
    const i64 e = min(begin + len, s + chunk);
    double acc = 0.0;
    for (i64 i = s + threadIdx.x; i < e; i += DOT_THREADS) acc += x[i] * y[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = DOT_THREADS / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(DOT_THREADS) void dot_final_kernel(const double* __restrict__ part, int G,
                                                                 double* __restrict__ out) {
    __shared__ double red[DOT_THREADS];
    double acc = 0.0;
    for (int i = threadIdx.x; i < G; i += DOT_THREADS) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = DOT_THREADS / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

int hs_launch_dot(hs_handle_t h, const double* x, const double* y, i64 begin, i64 len, double* part, int maxG,
                  double* out) {
    i64 G = (len + 4095) / 4096;
    G = std::max<i64>(1, std::min<i64>(G, maxG));
    const i64 chunk = (len + G - 1) / G;
    dot_partial_kernel<<<dim3((unsigned)G), dim3(DOT_THREADS), 0, h->stream>>>(x, y, begin, len, chunk, part);
    HS_LAUNCH_CHECK();
    dot_final_kernel<<<1, DOT_THREADS, 0, h->stream>>>(part, (int)G, out);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// ---------------------------------------------------------------------------
// C-ABI entries
// ---------------------------------------------------------------------------
extern "C" int hs_potrf_diag_f64(hs_handle_t h, double* A, int64_t N, int b, int blockID, int blockRow) {
    if (!h) return -1;
    if (!A) return -2;
    if (N < 0) return -3;
    if (b <= 0) return -4;
    if (blockID < 0) return -5;
    if (blockRow < 0) return -6;
    return hs_launch_potrf(h, A, N, b, blockID, blockRow);
}

extern "C" int hs_trsm_panel_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                                 int blockCount) {
    if (!h) return -1;
    if (!A) return -2;
    if (b <= 0) return -3;
    if (blockID < 0) return -4;
    if (blockRow < 0) return -5;
    if (blockStart < 0) return -6;
    if (blockCount < 0) return -7;
    return hs_launch_trsm(h, A, b, blockID, blockRow, blockStart, blockCount);
}

extern "C" int hs_syrk_diag_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                                int blockCount, int blockCountXY) {
    if (!h) return -1;
    if (!A) return -2;
    if (b <= 0) return -3;
    if (blockID != (int)hs_block_id(blockRow, blockRow, blockCountXY)) return -4;
    if (blockRow < 0 || blockRow >= blockCountXY) return -5;
    if (blockStart <= blockRow) return -6;
    if (blockCount < 0 || blockStart + blockCount > blockCountXY) return -7;
    return hs_launch_trailing(h, A, b, blockRow, blockStart, blockCount, blockCountXY, 1);
}

extern "C" int hs_gemm_trailing_f64(hs_handle_t h, double* A, int b, int blockID, int blockRow, int blockStart,
                                    int blockCount, int blockCountXY) {
    if (!h) return -1;
    if (!A) return -2;
    if (b <= 0) return -3;
    if (blockID != (int)hs_block_id(blockRow, blockRow, blockCountXY)) return -4;
    if (blockRow < 0 || blockRow >= blockCountXY) return -5;
    if (blockStart < blockRow + 2) return -6;
    if (blockCount < 0 || blockStart + blockCount > blockCountXY) return -7;
    return hs_launch_trailing(h, A, b, blockRow, blockStart, blockCount, blockCountXY, 2);
}

extern "C" int hs_syrk_gemm_trailing_f64(hs_handle_t h, double* A, int b, int k, int blockCountXY) {
    if (!h) return -1;
    if (!A) return -2;
    if (b <= 0) return -3;
    if (k < 0 || k >= blockCountXY) return -4;
    return hs_launch_trailing(h, A, b, k, k + 1, blockCountXY - k - 1, blockCountXY, 3);
}

extern "C" int hs_symv_block_f64(hs_handle_t h, const double* A, const double* x, double* result, int b,
                                 int blockStart_i, int blockStart_j, int blockCount_i, int blockCount_j,
                                 int blockCountXY, int reset) {
    if (!h) return -1;
    if (!A) return -2;
    if (!x) return -3;
    if (!result) return -4;
    if (b <= 0) return -5;
    if (blockStart_i < 0 || blockStart_i > blockCountXY) return -6;
    if (blockStart_j < 0 || blockStart_j > blockCountXY) return -7;
    if (blockCount_i < 0 || blockStart_i + blockCount_i > blockCountXY) return -8;
    if (blockCount_j < 0 || blockStart_j + blockCount_j > blockCountXY) return -9;
    if (blockCount_i == 0) return HS_OK;
    const i64 rows = (i64)blockCount_i * b;
    symv_rows_kernel<<<dim3((unsigned)((rows + 127) / 128)), dim3(128), 0, h->stream>>>(
        A, x, result, b, blockStart_i, blockStart_j, blockCount_i, blockCount_j, blockCountXY, reset);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

extern "C" int hs_trsv_block_f64(hs_handle_t h, const double* A, double* x, int64_t N, int b, int blockRow,
                                 int blockID, int transposed) {
    if (!h) return -1;
    if (!A) return -2;
    if (!x) return -3;
    if (N < 0) return -4;
    if (b <= 0) return -5;
    if (blockRow < 0) return -6;
    if (blockID < 0) return -7;
    return hs_launch_trsv_block(h, A, x, N, b, blockRow, blockID, transposed);
}

extern "C" int hs_col_update_f64(hs_handle_t h, const double* A, double* x, int b, int blockStart, int blockCount,
                                 int blockRow, int blockID, int blockCountXY, int transposed) {
    if (!h) return -1;
    if (!A) return -2;
    if (!x) return -3;
    if (b <= 0) return -4;
    if (blockStart < 0) return -5;
    if (blockCount < 0 || blockStart + blockCount > blockCountXY) return -6;
    if (blockRow < 0 || blockRow >= blockCountXY) return -7;
    if (blockID < 0) return -8;
    return hs_launch_col_update(h, A, x, b, blockStart, blockCount, blockRow, blockID, blockCountXY, transposed);
}

extern "C" int hs_vec_scale_f64(hs_handle_t h, const double* x, double alpha, double* result, int b, int blockStart,
                                int blockCount) {
    return launch_vec(h, x, nullptr, alpha, result, b, blockStart, blockCount, 0);
}
extern "C" int hs_vec_add_f64(hs_handle_t h, const double* x, const double* y, double* result, int b, int blockStart,
                              int blockCount) {
    return launch_vec(h, x, y, 0.0, result, b, blockStart, blockCount, 1);
}
extern "C" int hs_vec_sub_f64(hs_handle_t h, const double* x, const double* y, double* result, int b, int blockStart,
                              int blockCount) {
    return launch_vec(h, x, y, 0.0, result, b, blockStart, blockCount, 2);
}
extern "C" int hs_vec_scale_add_f64(hs_handle_t h, const double* x, double alpha, const double* y, double* result,
                                    int b, int blockStart, int blockCount) {
    return launch_vec(h, x, y, alpha, result, b, blockStart, blockCount, 3);
}

extern "C" int hs_read_sum_f64(hs_handle_t h, const double* x, int64_t count, double* partials_dev) {
    if (!h) return -1;
    if (!x) return -2;
    if (count <= 0) return -3;
    if (!partials_dev) return -4;
    const i64 groups = (count + RS_CHUNK - 1) / RS_CHUNK;
    if (groups > 0x7fffffffLL) {
        hs_set_error("hs_read_sum_f64: count too large");
        return -3;
    }
    const int vec = ((uintptr_t)x % 16 == 0) ? 1 : 0;
    read_sum_kernel<<<dim3((unsigned)groups), dim3(256), 0, h->stream>>>(x, count, partials_dev, vec);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

size_t hs_ws_dot() { return sizeof(double) * 1024; }

extern "C" int hs_dot_f64(hs_handle_t h, const double* x, const double* y, int b, int blockStart, int blockCount,
                          double* result_dev) {
    if (!h) return -1;
    if (!x) return -2;
    if (!y) return -3;
    if (b <= 0) return -4;
    if (blockStart < 0) return -5;
    if (blockCount < 0) return -6;
    if (!result_dev) return -7;
    const int maxG = 1024;
    double* part = nullptr;
    int rc = hs_work(h, hs_ws_dot(), (void**)&part);
    if (rc) return rc;
    return hs_launch_dot(h, x, y, (i64)blockStart * b, (i64)blockCount * b, part, maxG, result_dev);
}

