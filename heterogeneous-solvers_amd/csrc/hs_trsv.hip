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

// ROWS rows per wave: 16 for the per-block steps; 2 for the 1024-row
// super-block steps, whose 16 loads per lane and row already keep a wave's
// loads in flight and which then spread over 128 work-groups.
template <typename T, int BS, int SUB, int ROWS = 16>
__global__ __launch_bounds__(TS_THREADS) void gemv_rows_kernel(const T* __restrict__ base, const double* __restrict__ vin,
                                                               double* __restrict__ out, int strips, i64 rowLimit) {
    constexpr int NL = BS / 128 > 0 ? BS / 128 : 1;  // 16-byte loads per lane per row
    constexpr int LPR = BS / 2 < 64 ? BS / 2 : 64;    // lanes per row
    constexpr int RPL = 64 / LPR;                      // rows per wave load
    constexpr int R = ROWS / RPL;                      // row sums per lane
    __shared__ double vs[BS];
    const int g = blockIdx.x / strips;
    const int r0 = (blockIdx.x % strips) * (4 * ROWS);
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

// ---- super-block steps ------------------------------------------------------
// Every dependent launch of the block substitution costs a launch-to-launch
// gap of several microseconds, which dominates the solves at n = 256 blocks.
// The solves therefore step over super-blocks of SB = 1024 rows (k = SB / b
// blocks): the inverses E_S = L_SS^{-1} of the diagonal super-blocks are
// formed once per factor (block forward substitution on the block inverses,
// E_ij = -L_ii^{-1} sum_{m=j}^{i-1} L_im E_mj, one level i - j at a time), and
// a step is then one SB x SB mat-vec plus one update over the k block
// columns: 2n / k dependent launches per solve instead of 2n.
constexpr int SB = 1024;

typedef double d4 __attribute__((ext_vector_type(4)));

// 64 x 64 tile of C += A B (A row-major of any stored type, B row-major
// double) on v_mfma_f64_16x16x4f64: 256 threads as 2 x 2 waves of 32 x 32,
// 16-deep k stages through LDS (A stored k-major), the next stage's loads in
// registers while the current one is multiplied. acc[i][j] register v of
// lane l is row 32 wm + 16 i + l/16 + 4v, column 32 wn + 16 j + l%16.
template <typename TA>
__device__ inline void gemm64_mfma(const TA* A, i64 lda, const double* B, i64 ldb, int K, d4 (&acc)[2][2],
                                   double (*As)[68], double (*Bs)[68]) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int m0 = (wave >> 1) * 32, n0 = (wave & 1) * 32;
    const int lr = lane & 15, lq = lane >> 4;
    double ra[4], rb[4];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = t + 256 * e;  // 0 .. 1023
            ra[e] = ld(A + (i64)(idx >> 4) * lda + k0 + (idx & 15));
            rb[e] = B[(i64)(k0 + (idx >> 6)) * ldb + (idx & 63)];
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int idx = t + 256 * e;
            As[idx & 15][idx >> 4] = ra[e];
            Bs[idx >> 6][idx & 63] = rb[e];
        }
        __syncthreads();
        if (k0 + 16 < K) fetch(k0 + 16);
#pragma unroll
        for (int ks = 0; ks < 16; ks += 4) {
            double av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = As[ks + lq][m0 + 16 * i + lr];
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = Bs[ks + lq][n0 + 16 * j + lr];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
}

// out[r][c] = sg * acc for the 64 x 64 tile (row pitch ldo)
__device__ inline void store64(double* out, i64 ldo, const d4 (&acc)[2][2], double sg) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = (wave >> 1) * 32, n0 = (wave & 1) * 32;
    const int lr = lane & 15, lq = lane >> 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) out[(i64)(m0 + 16 * i + lq + 4 * v) * ldo + n0 + 16 * j + lr] = sg * acc[i][j][v];
}

// E_S <- the block inverses on its diagonal, zero elsewhere
__global__ void super_init_kernel(const double* __restrict__ Dlo, int b, int k, i64 total, double* __restrict__ E) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const i64 S = e / ((i64)SB * SB);
    const int r = (int)((e / SB) % SB), c = (int)(e % SB);
    const int rb = r / b, cb = c / b;
    E[e] = (rb == cb) ? Dlo[((i64)S * k + rb) * b * b + (i64)(r % b) * b + (c % b)] : 0.0;
}

// level d, first half: Tb(S, j) = sum_{m=j}^{j+d-1} L(Sk+j+d, Sk+m) E_S(m, j)
template <typename T>
__global__ __launch_bounds__(256) void super_level_t_kernel(const T* __restrict__ L, const double* __restrict__ E,
                                                            double* __restrict__ Tb, int b, int n, int k, int d) {
    __shared__ double As[16][68], Bs[16][68];
    const int tpe = b / 64, tiles = tpe * tpe;
    const int batch = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const int S = batch / (k - d), j = batch % (k - d), i = j + d;
    const int tm = tile / tpe, tn = tile % tpe;
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = d4{0.0, 0.0, 0.0, 0.0};
    const double* ES = E + (i64)S * SB * SB;
    for (int m = j; m < i; ++m) {
        const T* A = L + hs_block_id(S * k + i, S * k + m, n) * (i64)b * b + (i64)tm * 64 * b;
        const double* B = ES + (i64)(m * b) * SB + j * b + tn * 64;
        gemm64_mfma(A, b, B, SB, b, acc, As, Bs);
    }
    store64(Tb + ((i64)S * k + j) * b * b + (i64)tm * 64 * b + tn * 64, b, acc, 1.0);
}

// level d, second half: E_S(j+d, j) = -L(Sk+j+d)^{-1} Tb(S, j)
__global__ __launch_bounds__(256) void super_level_e_kernel(const double* __restrict__ Dlo, const double* __restrict__ Tb,
                                                            double* __restrict__ E, int b, int k, int d) {
    __shared__ double As[16][68], Bs[16][68];
    const int tpe = b / 64, tiles = tpe * tpe;
    const int batch = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const int S = batch / (k - d), j = batch % (k - d), i = j + d;
    const int tm = tile / tpe, tn = tile % tpe;
    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = d4{0.0, 0.0, 0.0, 0.0};
    const double* A = Dlo + ((i64)S * k + i) * b * b + (i64)tm * 64 * b;
    const double* B = Tb + ((i64)S * k + j) * b * b + tn * 64;
    gemm64_mfma(A, b, B, b, b, acc, As, Bs);
    store64(E + (i64)S * SB * SB + (i64)(i * b + tm * 64) * SB + j * b + tn * 64, SB, acc, -1.0);
}

// Et_S = E_S^T through 32 x 32 LDS tiles
__global__ __launch_bounds__(256) void super_transpose_kernel(const double* __restrict__ E, double* __restrict__ Et) {
    __shared__ double tl[32][33];
    constexpr int TPE = SB / 32;
    const i64 S = blockIdx.x / (TPE * TPE);
    const int tile = blockIdx.x % (TPE * TPE);
    const int tr = tile / TPE, tc = tile % TPE;
    const double* src = E + S * SB * SB;
    double* dst = Et + S * SB * SB;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) tl[r][tx] = src[(i64)(tr * 32 + r) * SB + tc * 32 + tx];
    __syncthreads();
    for (int r = ty; r < 32; r += 8) dst[(i64)(tc * 32 + r) * SB + tr * 32 + tx] = tl[tx][r];
}

// Two adjacent stored values, loaded as one 16-byte (double) or 8-byte
// (float) access.
template <typename T>
using pair_t = T __attribute__((ext_vector_type(2)));

// Forward super-step update: y_I -= sum_{c<K} L(I, j0+c) z_{j0+c} for the
// block rows I = I0 + g. Work-group (g, strip) owns ROWS rows of block row I;
// wave w takes the blocks c = w, w+4, ... and stages all of a block's loads
// in registers before its products, so that they are in flight together
// (written as one accumulation loop the compiler waits on each load in
// turn). The four wave sums are added in wave order. The K blocks' z values
// sit in LDS.
template <typename T, int BS, int K, int ROWS>
__global__ __launch_bounds__(TS_THREADS) void gemv_rows_multi_kernel(const T* __restrict__ L,
                                                                     const double* __restrict__ zs,
                                                                     double* __restrict__ y, int n, int I0, int j0,
                                                                     int strips) {
    constexpr int NL = BS / 128 > 0 ? BS / 128 : 1;
    constexpr int LPR = BS / 2 < 64 ? BS / 2 : 64;
    constexpr int RPL = 64 / LPR;
    constexpr int R = ROWS / RPL;
    static_assert(K % 4 == 0, "four waves share the blocks evenly");
    __shared__ double vs[K * BS];
    __shared__ double red[4][ROWS];
    const int g = blockIdx.x / strips;
    const int I = I0 + g;
    const int r0 = (blockIdx.x % strips) * ROWS;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int i = t; i < K * BS; i += TS_THREADS) vs[i] = zs[i];
    __syncthreads();
    const int sub = lane / LPR, sl = lane % LPR;
    double p[R];
#pragma unroll
    for (int q = 0; q < R; ++q) p[q] = 0.0;
    for (int c = wave; c < K; c += 4) {
        const T* M = L + hs_block_id(I, j0 + c, n) * (i64)BS * BS + (i64)(r0 + sub) * BS + 2 * sl;
        pair_t<T> raw[R][NL];
#pragma unroll
        for (int kk = 0; kk < R; ++kk)
#pragma unroll
            for (int q = 0; q < NL; ++q) raw[kk][q] = *(const pair_t<T>*)(M + (i64)kk * RPL * BS + q * 128);
        const double* v = vs + c * BS + 2 * sl;
#pragma unroll
        for (int kk = 0; kk < R; ++kk) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < NL; ++q)
                acc += (double)raw[kk][q].x * v[q * 128] + (double)raw[kk][q].y * v[q * 128 + 1];
            p[kk] += acc;
        }
    }
    const double s = halving_reduce16<R, LPR>(p);
    constexpr int SH = __builtin_ctz(LPR / R);
    if ((sl & (LPR / R - 1)) == 0) red[wave][(sl >> SH) * RPL + sub] = s;
    __syncthreads();
    if (t < ROWS) y[(i64)I * BS + r0 + t] -= (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
}

// rows per work-group of the forward super-step update
template <int BS>
constexpr int multi_rows() { return BS == 64 ? 16 : 8; }

// Backward super-step update: z_I -= sum_{c<K} L(j0+c, I)^T x_{j0+c} for
// I < j0. Work-group (I, 64-column slice): lane l reads columns 2(l%32),
// 2(l%32)+1 of every other row (l/32 picks the row parity) and wave w a
// 1/W share of each block's rows, in staged chunks of up to 32 rows whose
// loads are all issued before their products. The 2W partial column sums are
// added in a fixed order. W = 16 for the short steps near the top, where
// few work-groups each stream K blocks.
template <typename T, int BS, int K, int W>
__global__ __launch_bounds__(64 * W) void gemv_cols_multi_kernel(const T* __restrict__ L,
                                                                 const double* __restrict__ xs_g,
                                                                 double* __restrict__ z, int n, int j0) {
    constexpr int CW = 64;
    constexpr int SL = BS / CW;
    constexpr int RQ = BS / W;               // rows per wave and block
    constexpr int CH = RQ < 32 ? RQ : 32;    // rows per staged chunk
    __shared__ double xs[K * BS];
    __shared__ double red[2 * W][CW];
    const int I = blockIdx.x / SL;
    const int c0 = (blockIdx.x % SL) * CW;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int par = lane >> 5, cl = lane & 31;
    for (int i = t; i < K * BS; i += 64 * W) xs[i] = xs_g[i];
    __syncthreads();
    double acc0 = 0.0, acc1 = 0.0;
    for (int c = 0; c < K; ++c) {
        const T* base = L + hs_block_id(j0 + c, I, n) * (i64)BS * BS + c0 + 2 * cl;
        const double* x = xs + c * BS + par;
#pragma unroll
        for (int r0 = wave * RQ; r0 < wave * RQ + RQ; r0 += CH) {
            pair_t<T> raw[CH / 2];
#pragma unroll
            for (int k = 0; k < CH / 2; ++k) raw[k] = *(const pair_t<T>*)(base + (i64)(r0 + 2 * k + par) * BS);
#pragma unroll
            for (int k = 0; k < CH / 2; ++k) {
                const double xr = x[r0 + 2 * k];
                acc0 += (double)raw[k].x * xr;
                acc1 += (double)raw[k].y * xr;
            }
        }
    }
    red[2 * wave + par][2 * cl] = acc0;
    red[2 * wave + par][2 * cl + 1] = acc1;
    __syncthreads();
    if (t < CW) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 2 * W; q += 4) s += (red[q][t] + red[q + 1][t]) + (red[q + 2][t] + red[q + 3][t]);
        z[(i64)I * BS + c0 + t] -= s;
    }
}

inline int super_k(int b) { return SB / b; }
inline int super_count(int n, int b) {
    const int k = super_k(b);
    return (k > 1 && n >= 2 * k) ? n / k : 0;
}

}  // namespace

bool hs_trsv_fast_ok(int b) { return b == 64 || b == 128 || b == 256; }

i64 hs_dinv_doubles(int n, int b) {
    const i64 base = 2 * (i64)n * b * b;
    if (!hs_trsv_fast_ok(b)) return base;
    return base + 2 * (i64)super_count(n, b) * SB * SB;
}

// Dinv (hs_dinv_doubles(n, b) doubles) holds two arrays of n b x b blocks,
// D_j = L_jj^{-T} then L_jj^{-1}, followed for the fast block sizes by the
// super-block inverses E_S (SB x SB, row-major) and their transposes.
template <typename T>
int hs_diag_inverse(hs_handle_t h, const T* L, i64 N, int b, double* Dinv) {
    const int n = (int)((N + b - 1) / b);
    const int strips = b / 16;
    const double* Dlo = Dinv + (i64)n * b * b;
    diag_inverse_kernel<T><<<dim3((unsigned)(n * strips)), dim3(TS_THREADS), 0, h->stream>>>(
        L, b, n, N, Dinv, Dinv + (i64)n * b * b, strips);
    HS_LAUNCH_CHECK();
    const int ns = hs_trsv_fast_ok(b) ? super_count(n, b) : 0;
    if (ns == 0) return HS_OK;
    const int k = super_k(b);
    double* E = Dinv + 2 * (i64)n * b * b;
    double* Et = E + (i64)ns * SB * SB;  // level scratch until the transpose
    const i64 total = (i64)ns * SB * SB;
    super_init_kernel<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, h->stream>>>(Dlo, b, k, total, E);
    HS_LAUNCH_CHECK();
    const int tiles = (b / 64) * (b / 64);
    for (int d = 1; d < k; ++d) {
        const unsigned grid = (unsigned)((i64)ns * (k - d) * tiles);
        super_level_t_kernel<T><<<dim3(grid), dim3(256), 0, h->stream>>>(L, E, Et, b, n, k, d);
        HS_LAUNCH_CHECK();
        super_level_e_kernel<<<dim3(grid), dim3(256), 0, h->stream>>>(Dlo, Et, E, b, k, d);
        HS_LAUNCH_CHECK();
    }
    super_transpose_kernel<<<dim3((unsigned)((i64)ns * (SB / 32) * (SB / 32))), dim3(256), 0, h->stream>>>(E, Et);
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
    constexpr int K = SB / BS;                     // blocks per super-block
    const int ns = super_count(n, BS);
    const double* E = Dinv + 2 * (i64)n * BS * BS;  // L_SS^{-1}
    const double* Et = E + (i64)ns * SB * SB;       // L_SS^{-T}
    const int jr = ns * K;                          // first block row past the super-blocks
    const dim3 blk(TS_THREADS);
    if (phases & 1) {
    int rc = hs_memset_zero(h, z, sizeof(double) * (size_t)n * BS);
    if (rc) return rc;
    for (int S = 0; S < ns; ++S) {
        // z_S = L_SS^{-1} y_S, then y_I -= sum_c L_I,c z_c below the super-block
        gemv_rows_kernel<double, SB, 0, 2><<<dim3(SB / 8), blk, 0, h->stream>>>(
            E + (i64)S * SB * SB, v + (i64)S * SB, z + (i64)S * SB, SB / 8, N - (i64)S * SB);
        HS_LAUNCH_CHECK();
        const int I0 = (S + 1) * K;
        if (I0 < n) {
            constexpr int MR = multi_rows<BS>();
            gemv_rows_multi_kernel<T, BS, K, MR><<<dim3((unsigned)((n - I0) * (BS / MR))), blk, 0, h->stream>>>(
                L, z + (i64)S * SB, v, n, I0, S * K, BS / MR);
            HS_LAUNCH_CHECK();
        }
    }
    for (int j = jr; j < n; ++j) {
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
    for (int j = n - 1; j >= jr; --j) {
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
    for (int S = ns - 1; S >= 0; --S) {
        // x_S = L_SS^{-T} z_S, then z_I -= sum_c L_c,I^T x_c above the super-block
        gemv_rows_kernel<double, SB, 0, 2><<<dim3(SB / 8), blk, 0, h->stream>>>(
            Et + (i64)S * SB * SB, z + (i64)S * SB, v + (i64)S * SB, SB / 8, N - (i64)S * SB);
        HS_LAUNCH_CHECK();
        if (S > 0) {
            const unsigned wgs = (unsigned)(S * K * (BS / 64));
            if (wgs < 2u * (unsigned)h->cu_count)
                gemv_cols_multi_kernel<T, BS, K, 16><<<dim3(wgs), dim3(1024), 0, h->stream>>>(
                    L, v + (i64)S * SB, z, n, S * K);
            else
                gemv_cols_multi_kernel<T, BS, K, 4><<<dim3(wgs), blk, 0, h->stream>>>(L, v + (i64)S * SB, z, n, S * K);
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
