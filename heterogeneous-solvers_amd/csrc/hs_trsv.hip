// Blocked triangular solves L z = y and L^T x = z on the packed factor, for
// the solver block sizes (64, 128, 256), with the factor stored in FP64 or
// FP32 (the mixed-precision factor); arithmetic is FP64 throughout.
//
// The reference (TriangularSystemSolver.cpp:31-82) runs, per block column j,
// a one-work-group TRSV with b serial barrier steps and then the column
// update. Here the diagonal blocks are inverted once (D_j = L_jj^{-T}, upper
// triangular, computed with the panel TRSM algorithm on identity blocks), so
// each step's diagonal solve is a b x b mat-vec, and the off-diagonal
// updates stream each stored block once per solve:
//   forward  step j:  z_j = D_j^T y_j ;  y_I -= L_Ij z_j      (I > j, column j)
//   backward step j:  x_j = D_j z_j   ;  z_I -= L_jI^T x_j    (I < j, row j)
#include <algorithm>

#include "hs_internal.h"

namespace {

constexpr int TS_THREADS = 256;

template <typename T>
__device__ inline double ld(const T* p) {
    return (double)*p;
}

// D_j = L_jj^{-T}: work-group (j, strip) solves X L^T = I for 16 rows of X,
// thread t owning column t (same recurrence as the panel TRSM). Rows and
// columns at or past N (padding, zero diagonal) stay zero in D, which makes
// the solves leave padded entries untouched, as the reference's TRSV does
// (MatrixVectorOperations.cpp:314-320).
template <typename T>
__global__ __launch_bounds__(TS_THREADS) void diag_inverse_kernel(const T* __restrict__ L, int b, int n, i64 N,
                                                                  double* __restrict__ Dinv, double* __restrict__ Linv,
                                                                  int strips) {
    constexpr int TR = 16;
    __shared__ double xs[2][TR];
    const int j = blockIdx.x / strips;
    const int r0 = (blockIdx.x % strips) * TR;
    const int t = threadIdx.x;
    const int nvalid = (int)min((i64)b, N - (i64)j * b);
    const bool active = t < nvalid;
    const T* Lj = L + hs_block_id(j, j, n) * (i64)b * b;
    const T* Lrow = Lj + (i64)(active ? t : 0) * b;
    double x[TR];
#pragma unroll
    for (int i = 0; i < TR; ++i) x[i] = (active && r0 + i == t) ? 1.0 : 0.0;
    for (int c = 0; c < nvalid; ++c) {
        const int buf = c & 1;
        if (t == c) {
            const double lcc = ld(Lrow + c);
#pragma unroll
            for (int i = 0; i < TR; ++i) {
                x[i] = x[i] / lcc;
                xs[buf][i] = x[i];
            }
        }
        __syncthreads();
        if (t > c && active) {
            const double ltc = ld(Lrow + c);
#pragma unroll
            for (int i = 0; i < TR; ++i) x[i] = x[i] - ltc * xs[buf][i];
        }
    }
    if (t < b) {
        double* D = Dinv + (i64)j * b * b;
        double* Li = Linv + (i64)j * b * b + (i64)t * b + r0;  // row t of D^T
#pragma unroll
        for (int i = 0; i < TR; ++i) {
            const double v = (r0 + i < nvalid) ? x[i] : 0.0;
            D[(i64)(r0 + i) * b + t] = v;
            Li[i] = v;
        }
    }
}

// Row-dot block GEMV, the workhorse of both solves: for consecutive b x b
// blocks M_g = base + g*b*b (g < gridDim.x / strips), work-group (g, strip)
// computes 64 rows r of  s_r = sum_c M_g[r][c] v_in[c]  and then
//   SUB = 1:  out_g[r] -= s_r          (off-diagonal update)
//   SUB = 0:  out_g[r]  = s_r, r < N   (diagonal step, in place: v_in = out_0)
// Wave w owns rows 16w..16w+15 of the strip; lane l reads 16 contiguous bytes
// (columns 2l, 2l+1, then +128 for b = 256) of every row, so each wave keeps
// 16-32 independent loads in flight; the 16 row sums are finished with a
// recursive-halving shuffle reduction.
template <int R, int G>
__device__ inline double halving_reduce16(double (&v)[R]) {
    const int lane = threadIdx.x & 63;
    int len = R;
#pragma unroll
    for (int m = G / 2; m >= G / R; m >>= 1) {
        const int half = len >> 1;
        const bool upper = (lane & m) != 0;
#pragma unroll
        for (int i = 0; i < R / 2; ++i) {
            if (i < half) {
                const double send = upper ? v[i] : v[i + half];
                const double keep = upper ? v[i + half] : v[i];
                v[i] = keep + __shfl_xor(send, m, 64);
            }
        }
        len = half;
    }
    double s = v[0];
#pragma unroll
    for (int m = G / R / 2; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    return s;
}

template <typename T, int BS, int SUB>
__global__ __launch_bounds__(TS_THREADS) void gemv_rows_kernel(const T* __restrict__ base, const double* __restrict__ vin,
                                                               double* __restrict__ out, int strips, i64 rowLimit) {
    constexpr int NL = BS / 128 > 0 ? BS / 128 : 1;  // 16-byte loads per lane per row
    constexpr int LPR = BS / 2 < 64 ? BS / 2 : 64;    // lanes per row
    constexpr int RPL = 64 / LPR;                      // rows per wave load
    constexpr int ROWS = 16;                           // rows per wave
    constexpr int R = ROWS / RPL;                      // row sums per lane
    __shared__ double vs[BS];
    const int g = blockIdx.x / strips;
    const int r0 = (blockIdx.x % strips) * 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int i = t; i < BS; i += TS_THREADS) vs[i] = vin[i];
    __syncthreads();
    const T* M = base + (i64)g * BS * BS;
    const int sub = lane / LPR, sl = lane % LPR;
    double p[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int r = r0 + wave * ROWS + k * RPL + sub;
        const T* row = M + (i64)r * BS + 2 * sl;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            acc += ld(row + q * 128) * vs[q * 128 + 2 * sl] + ld(row + q * 128 + 1) * vs[q * 128 + 2 * sl + 1];
        }
        p[k] = acc;
    }
    const double s = halving_reduce16<R, LPR>(p);
    constexpr int SH = __builtin_ctz(LPR / R);
    if ((sl & (LPR / R - 1)) == 0) {
        const int r = r0 + wave * ROWS + (sl >> SH) * RPL + sub;
        double* o = out + (i64)g * BS + r;
        if (SUB) *o -= s;
        else if (r < rowLimit) *o = s;
    }
}

// Column-dot update of the backward solve: one work-group per block (j, I),
// I < j (row j of the block triangle): z_I -= L_jI^T x_j. Lane l owns columns
// 2l, 2l+1 (and +128 for b = 256); wave w sums a quarter of the rows with 8
// rows of loads in flight, the quarters are added in wave order.
template <typename T, int BS>
__global__ __launch_bounds__(TS_THREADS) void gemv_cols_update_kernel(const T* __restrict__ L,
                                                                      const double* __restrict__ v,
                                                                      double* __restrict__ z, int n, int j) {
    constexpr int NL = BS / 128 > 0 ? BS / 128 : 1;
    constexpr int LPR = BS / 2 < 64 ? BS / 2 : 64;
    constexpr int RPL = 64 / LPR;
    constexpr int RQ = BS / 4;  // rows per wave
    __shared__ double xs[BS];
    __shared__ double red[4 * RPL][BS];
    const int I = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int i = t; i < BS; i += TS_THREADS) xs[i] = v[(i64)j * BS + i];
    __syncthreads();
    const T* M = L + hs_block_id(j, I, n) * (i64)BS * BS;
    const int sub = lane / LPR, sl = lane % LPR;
    double acc[2 * NL];
#pragma unroll
    for (int q = 0; q < 2 * NL; ++q) acc[q] = 0.0;
#pragma unroll 8
    for (int k = 0; k < RQ / RPL; ++k) {
        const int r = wave * RQ + k * RPL + sub;
        const T* row = M + (i64)r * BS + 2 * sl;
        const double xr = xs[r];
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            acc[2 * q] += ld(row + q * 128) * xr;
            acc[2 * q + 1] += ld(row + q * 128 + 1) * xr;
        }
    }
#pragma unroll
    for (int q = 0; q < NL; ++q) {
        red[wave * RPL + sub][q * 128 + 2 * sl] = acc[2 * q];
        red[wave * RPL + sub][q * 128 + 2 * sl + 1] = acc[2 * q + 1];
    }
    __syncthreads();
    double* zI = z + (i64)I * BS;
    for (int c = t; c < BS; c += TS_THREADS) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < 4 * RPL; ++w) s += red[w][c];
        zI[c] -= s;
    }
}

// Column-slice form of the same update for the steps with few blocks: work-
// group (I, slice) owns CW = 64 columns of block (j, I), lane l column
// slice * 64 + l, wave w a quarter of the rows (one 8-byte load per lane and
// row, eight rows in flight); the four wave sums are added in wave order.
// Four work-groups per block instead of one keep more CUs streaming while the
// block row j is short.
template <typename T, int BS>
__global__ __launch_bounds__(TS_THREADS) void gemv_cols_slice_kernel(const T* __restrict__ L,
                                                                     const double* __restrict__ v,
                                                                     double* __restrict__ z, int n, int j) {
    constexpr int CW = 64;
    constexpr int SL = BS / CW;  // slices per block
    constexpr int RQ = BS / 4;   // rows per wave
    __shared__ double xs[BS];
    __shared__ double red[4][CW];
    const int I = blockIdx.x / SL;
    const int c0 = (blockIdx.x % SL) * CW;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int i = t; i < BS; i += TS_THREADS) xs[i] = v[(i64)j * BS + i];
    __syncthreads();
    const T* col = L + hs_block_id(j, I, n) * (i64)BS * BS + c0 + lane;
    double acc = 0.0;
#pragma unroll 8
    for (int k = 0; k < RQ; ++k) {
        const int r = wave * RQ + k;
        acc += ld(col + (i64)r * BS) * xs[r];
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (t < CW) z[(i64)I * BS + c0 + t] -= (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
}

}  // namespace

bool hs_trsv_fast_ok(int b) { return b == 64 || b == 128 || b == 256; }

// Dinv holds two arrays of n b x b blocks: D_j = L_jj^{-T} then L_jj^{-1}.
template <typename T>
int hs_diag_inverse(hs_handle_t h, const T* L, i64 N, int b, double* Dinv) {
    const int n = (int)((N + b - 1) / b);
    const int strips = b / 16;
    diag_inverse_kernel<T><<<dim3((unsigned)(n * strips)), dim3(TS_THREADS), 0, h->stream>>>(
        L, b, n, N, Dinv, Dinv + (i64)n * b * b, strips);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// v holds y on entry and x on exit; z (n*b, scratch) receives the forward
// solution. Each diagonal step reads one vector and writes the other, so the
// strips of one launch never read what another strip writes. phases: bit 0
// runs the forward solve (z = L^{-1} v, v clobbered), bit 1 the backward
// solve (v = L^{-T} z, z clobbered).
template <typename T, int BS>
static int tri_solve_bs(hs_handle_t h, const T* L, const double* Dinv, double* v, double* z, i64 N, int phases) {
    const int n = (int)((N + BS - 1) / BS);
    const int strips = BS / 64;
    const double* Dup = Dinv;                      // D_j = L_jj^{-T}
    const double* Dlo = Dinv + (i64)n * BS * BS;   // L_jj^{-1}
    const dim3 blk(TS_THREADS);
    if (phases & 1) {
    int rc = hs_memset_zero(h, z, sizeof(double) * (size_t)n * BS);
    if (rc) return rc;
    for (int j = 0; j < n; ++j) {
        // z_j = L_jj^{-1} y_j (entries past N stay zero)
        gemv_rows_kernel<double, BS, 0><<<dim3((unsigned)strips), blk, 0, h->stream>>>(
            Dlo + (i64)j * BS * BS, v + (i64)j * BS, z + (i64)j * BS, strips, N - (i64)j * BS);
        HS_LAUNCH_CHECK();
        if (j + 1 < n) {
            // y_I -= L_Ij z_j over the contiguous blocks (j+1..n-1, j)
            gemv_rows_kernel<T, BS, 1><<<dim3((unsigned)((n - j - 1) * strips)), blk, 0, h->stream>>>(
                L + hs_block_id(j + 1, j, n) * BS * BS, z + (i64)j * BS, v + (i64)(j + 1) * BS, strips, 0);
            HS_LAUNCH_CHECK();
        }
    }
    }
    if (phases & 2) {
    for (int j = n - 1; j >= 0; --j) {
        // x_j = L_jj^{-T} z_j (entries past N keep their zero)
        gemv_rows_kernel<double, BS, 0><<<dim3((unsigned)strips), blk, 0, h->stream>>>(
            Dup + (i64)j * BS * BS, z + (i64)j * BS, v + (i64)j * BS, strips, N - (i64)j * BS);
        HS_LAUNCH_CHECK();
        if (j > 0) {
            // z_I -= L_jI^T x_j for blocks (j, I), I < j; column slices while
            // the row is shorter than the CU count
            if (BS >= 128 && j < h->cu_count)
                gemv_cols_slice_kernel<T, BS><<<dim3((unsigned)(j * (BS / 64))), blk, 0, h->stream>>>(L, v, z, n, j);
            else
                gemv_cols_update_kernel<T, BS><<<dim3((unsigned)j), blk, 0, h->stream>>>(L, v, z, n, j);
            HS_LAUNCH_CHECK();
        }
    }
    }
    return HS_OK;
}

// Both solves in place on v (n*b entries), factor L (packed, T) with the
// inverted diagonal blocks from hs_diag_inverse; z is an n*b scratch vector.
template <typename T>
int hs_tri_solve_prepared(hs_handle_t h, const T* L, const double* Dinv, double* v, double* z, i64 N, int b,
                          int phases) {
    switch (b) {
        case 64: return tri_solve_bs<T, 64>(h, L, Dinv, v, z, N, phases);
        case 128: return tri_solve_bs<T, 128>(h, L, Dinv, v, z, N, phases);
        case 256: return tri_solve_bs<T, 256>(h, L, Dinv, v, z, N, phases);
        default: hs_set_error("fast triangular solve needs b in {64, 128, 256}"); return HS_ERR_UNSUPPORTED;
    }
}

template int hs_diag_inverse<double>(hs_handle_t, const double*, i64, int, double*);
template int hs_diag_inverse<float>(hs_handle_t, const float*, i64, int, double*);
template int hs_tri_solve_prepared<double>(hs_handle_t, const double*, const double*, double*, double*, i64, int, int);
template int hs_tri_solve_prepared<float>(hs_handle_t, const float*, const double*, double*, double*, i64, int, int);
