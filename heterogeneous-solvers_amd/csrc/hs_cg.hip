// Symmetric packed mat-vec that reads every stored block once, and the
// device-resident CG loop built on it.
//
// SYMV (replaces matrixVectorBlock_GPU, MatrixVectorOperations.cpp:95-192,
// which reads every off-diagonal block twice, once per triangle):
//   pass 1, one work-group per stored block (I,J), in packed (memory) order:
//     rowpart[ID] = A_IJ x_J           (b values)
//     colpart[ID] = A_IJ^T x_I         (b values, I > J only)
//   pass 2, 64 rows of block row I per work-group, fixed order over J:
//     y_i = sum_{J<=I} rowpart[ID(I,J)]_i + sum_{J>I} colpart[ID(J,I)]_i
// Sixteen lanes share a row, each reading 16 contiguous bytes per load, so a
// wave load covers 256 bytes of each of 4 rows; row sums are finished by a
// 16-lane recursive-halving shuffle reduction, column sums stay in registers
// until an LDS combine.
// The result is bit-reproducible run to run (no atomics).
#include <algorithm>
#include <atomic>
#include <climits>
#include <mutex>

#include "hs_internal.h"

int hs_launch_dot(hs_handle_t h, const double* x, const double* y, i64 begin, i64 len, double* part, int maxG,
                  double* out);

namespace {

constexpr int SYMV_THREADS = 256;

// Recursive-halving reduction of R values per lane across aligned groups of
// G lanes. On return v[0] of lane l holds the group total of value index
// (l & (G-1)) >> log2(G/R).
template <int R, int G>
__device__ inline double halving_reduce(double (&v)[R]) {
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

__device__ inline void decode_block(i64 id, int n, int* I, int* J) {
    // column J starts at S(J) = J*n - J(J-1)/2
    const double nn = 2.0 * n + 1.0;
    i64 j = (i64)((nn - sqrt(nn * nn - 8.0 * (double)id)) * 0.5);
    if (j < 0) j = 0;
    while (j > 0 && j * n - j * (j - 1) / 2 > id) --j;
    while ((j + 1) * n - (j + 1) * j / 2 <= id) ++j;
    *J = (int)j;
    *I = (int)(j + (id - (j * n - j * (j - 1) / 2)));
}

// ABS: use |a_ij| (gives the row sums of |A| for x = 1, i.e. ||A||_inf)
// Layout: 16 lanes share a row; lane l owns the columns 2(l%16) + 32 s + {0,1},
// s < b/32, so one wave load covers 256 contiguous bytes of each of 4 rows
// (quarter q = l/16) and a wave walks b/4 rows in groups of 4. A row sum is
// then a 16-lane reduction (b/16 values over 16 lanes), which keeps the waves'
// time between their streaming loads short; the column partials of the 16
// (wave, quarter) pairs are added through LDS. Block data is read with
// non-temporal loads: each element is used once per pass. The partials are
// written with non-temporal stores too: written back through the cache they
// cost about 9% of the pass (measured at n = 256, b = 128).
template <int BS, bool ABS = false, int GB = 4, int ROLL = 1>
__global__ __launch_bounds__(SYMV_THREADS) void symv_partials_kernel(const double* __restrict__ A,
                                                                      const double* __restrict__ x,
                                                                      double* __restrict__ rowpart,
                                                                      double* __restrict__ colpart, int n,
                                                                      const i64* __restrict__ stop, i64 it) {
    if (stop && it > *stop) return;
    constexpr int NS = BS / 32;  // 16-byte loads per lane per row
    constexpr int RW = BS / 4;   // rows per wave
    constexpr int NG = RW / 4;   // 4-row groups per wave
    __shared__ double colred[16][BS];

    const i64 id = blockIdx.x;
    const double* blk = A + id * (i64)BS * BS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, c16 = lane & 15;
    // the block's (I, J) and the x segments are needed only for the products:
    // they are found after the first batch's block loads are issued, so that
    // the decode's square root and loops run under those loads
    int I = 0, J = 0;
    const double* xI = x;
    double xj[2 * NS], cacc[2 * NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        xj[2 * s] = xj[2 * s + 1] = 0.0;
        cacc[2 * s] = 0.0;
        cacc[2 * s + 1] = 0.0;
    }
    // GB row groups per batch: a batch's loads are issued together, then its
    // GB row sums are finished (16-lane reduction) and written, so only one
    // batch of loads and sums is live at a time (ROLL keeps the batch loop
    // rolled, which stops the compiler from hoisting later batches' loads).
    // Measured: b = 128 streams fastest with all 8 groups in one batch (2
    // waves per SIMD, 128 KB in flight per work-group); b = 256 with batches
    // of 2 (3 waves per SIMD; a whole block in registers spills).
#pragma unroll ROLL
    for (int g0 = 0; g0 < NG; g0 += GB) {
        double a[GB][2 * NS];
#pragma unroll
        for (int g = 0; g < GB; ++g) {
            const double* rp = blk + (i64)(wave * RW + 4 * (g0 + g) + q) * BS + 2 * c16;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                a[g][2 * s] = __builtin_nontemporal_load(rp + 32 * s);
                a[g][2 * s + 1] = __builtin_nontemporal_load(rp + 32 * s + 1);
            }
        }
        if (g0 == 0) {
            decode_block(id, n, &I, &J);
            xI = x + (i64)I * BS;
            const double* xJ = x + (i64)J * BS;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const double2 v = *reinterpret_cast<const double2*>(xJ + 2 * c16 + 32 * s);
                xj[2 * s] = v.x;
                xj[2 * s + 1] = v.y;
            }
        }
        double pb[GB];
#pragma unroll
        for (int g = 0; g < GB; ++g) {
            const double xi = xI[wave * RW + 4 * (g0 + g) + q];
            double acc = 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double a0 = a[g][2 * s], a1 = a[g][2 * s + 1];
                if (ABS) {
                    a0 = fabs(a0);
                    a1 = fabs(a1);
                }
                acc += a0 * xj[2 * s] + a1 * xj[2 * s + 1];
                cacc[2 * s] += a0 * xi;
                cacc[2 * s + 1] += a1 * xi;
            }
            pb[g] = acc;
        }
        constexpr int LPB = 16 / GB;  // lanes per row total after the reduction
        const double tot = halving_reduce<GB, 16>(pb);
        if ((c16 & (LPB - 1)) == 0)
            __builtin_nontemporal_store(tot, rowpart + id * BS + wave * RW + 4 * (g0 + c16 / LPB) + q);
    }
    if (I == J) return;  // diagonal blocks hold both triangles: row form only
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        colred[wave * 4 + q][2 * c16 + 32 * s] = cacc[2 * s];
        colred[wave * 4 + q][2 * c16 + 32 * s + 1] = cacc[2 * s + 1];
    }
    __syncthreads();
    for (int c = threadIdx.x; c < BS; c += SYMV_THREADS) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < 16; ++w) s += colred[w][c];
        __builtin_nontemporal_store(s, colpart + id * BS + c);
    }
}

// pass 2 variants: 0 = y = sum; 1 = y = sum and dqpart[wg] = sum_i y_i d_i;
// 2 = y = rhs - sum and part[wg] = sum_i y_i^2
// One work-group per 64 consecutive rows (b is a multiple of 64, so they lie
// in one block row I); wave w sums block columns [w n/4, (w+1) n/4) for its
// lane's row, the four wave sums are added in wave order.
constexpr int FIN_ROWS = 64;

template <int MODE>
__global__ __launch_bounds__(SYMV_THREADS) void symv_finish_kernel(const double* __restrict__ rowpart,
                                                                    const double* __restrict__ colpart, int n, int b,
                                                                    double* __restrict__ y,
                                                                    const double* __restrict__ other,
                                                                    double* __restrict__ part,
                                                                    const i64* __restrict__ stop, i64 it) {
    if (stop && it > *stop) return;
    __shared__ double red[4][FIN_ROWS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const i64 i = (i64)blockIdx.x * FIN_ROWS + lane;
    const int I = (int)(i / b), ii = (int)(i % b);
    const int j0 = (int)((i64)wave * n / 4), j1 = (int)((i64)(wave + 1) * n / 4);
    // 8 loads in flight per step; the sum itself stays in J order
    double s = 0.0;
    const int jr = min(j1, I + 1);
    int J = j0;
    for (; J + 8 <= jr; J += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = rowpart[hs_block_id(I, J + u, n) * b + ii];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; J < jr; ++J) s += rowpart[hs_block_id(I, J, n) * b + ii];
    const i64 base = hs_block_id(I, I, n);
    J = max(j0, I + 1);
    for (; J + 8 <= j1; J += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = colpart[(base + (J + u - I)) * b + ii];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; J < j1; ++J) s += colpart[(base + (J - I)) * b + ii];
    red[wave][lane] = s;
    __syncthreads();
    if (wave != 0) return;
    s = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    if (MODE == 2) s = other[i] - s;
    y[i] = s;
    if (MODE == 0) return;
    double v = (MODE == 1) ? s * other[i] : s * s;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if (lane == 0) part[blockIdx.x] = v;
}

// every work-group of the vector kernels sums the G partials in the same
// fixed tree order, so all of them see bit-identical scalars
__device__ inline double wg_sum_partials(const double* __restrict__ part, int G, double* red) {
    double a = 0.0;
    for (int g = threadIdx.x; g < G; g += SYMV_THREADS) a += part[g];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int st = SYMV_THREADS / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

struct CGState {
    double delta[2];  // delta_new of even / odd iteration parity
    double delta0;
    double eps2;
    double alpha;
    double pad;
    i64 stop;  // last executed iteration; LLONG_MAX while running
};

__global__ __launch_bounds__(SYMV_THREADS) void cg_init_kernel(const double* __restrict__ rhs, double* __restrict__ x,
                                                                double* __restrict__ r, double* __restrict__ d, i64 rows,
                                                                double* __restrict__ part) {
    __shared__ double red[SYMV_THREADS];
    const i64 i = (i64)blockIdx.x * SYMV_THREADS + threadIdx.x;
    double v = 0.0;
    if (i < rows) {
        // x0 = 0 so r = b - A x0 = b exactly (CG.cpp:277-306)
        const double bi = rhs[i];
        x[i] = 0.0;
        r[i] = bi;
        d[i] = bi;
        v = bi * bi;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int st = SYMV_THREADS / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(SYMV_THREADS) void cg_init_finish_kernel(const double* __restrict__ part, int G,
                                                                       CGState* st, double eps2) {
    __shared__ double red[SYMV_THREADS];
    const double s = wg_sum_partials(part, G, red);
    if (threadIdx.x == 0) {
        st->delta[0] = s;
        st->delta[1] = s;
        st->delta0 = s;
        st->eps2 = eps2;
        st->alpha = 0.0;
        st->stop = (s > eps2 * s) ? LLONG_MAX : -1;
    }
}

// alpha = delta / d.q ; x += alpha d ; (it % 50 != 0) r -= alpha q, part = r.r
__global__ __launch_bounds__(SYMV_THREADS) void cg_xr_kernel(CGState* st, const double* __restrict__ dqpart, int G,
                                                              const double* __restrict__ d,
                                                              const double* __restrict__ q, double* __restrict__ x,
                                                              double* __restrict__ r, i64 rows,
                                                              double* __restrict__ rrpart, i64 it, int updateR) {
    if (it > st->stop) return;
    __shared__ double red[SYMV_THREADS];
    const double dq = wg_sum_partials(dqpart, G, red);
    const double alpha = st->delta[it & 1] / dq;
    const i64 i = (i64)blockIdx.x * SYMV_THREADS + threadIdx.x;
    double v = 0.0;
    if (i < rows) {
        x[i] = x[i] + d[i] * alpha;
        if (updateR) {
            const double ri = r[i] + q[i] * (-1.0 * alpha);
            r[i] = ri;
            v = ri * ri;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = alpha;
    if (!updateR) return;
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = SYMV_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) rrpart[blockIdx.x] = red[0];
}

// delta_new = r.r ; beta = delta_new / delta_old ; d = r + beta d ; stop test
__global__ __launch_bounds__(SYMV_THREADS) void cg_d_kernel(CGState* st, const double* __restrict__ rrpart, int G,
                                                             const double* __restrict__ r, double* __restrict__ d,
                                                             i64 rows, i64 it) {
    if (it > st->stop) return;
    __shared__ double red[SYMV_THREADS];
    const double dn = wg_sum_partials(rrpart, G, red);
    const double beta = dn / st->delta[it & 1];
    const i64 i = (i64)blockIdx.x * SYMV_THREADS + threadIdx.x;
    if (i < rows) d[i] = r[i] + d[i] * beta;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->delta[(it + 1) & 1] = dn;
        if (!(dn > st->eps2 * st->delta0)) st->stop = it;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static bool symv_fast_ok(int b) { return b == 64 || b == 128 || b == 256; }

static int launch_partials_abs(hs_handle_t h, const double* A, const double* x, double* rowpart, double* colpart,
                              int n, int b) {
    const i64 blocks = (i64)n * (n + 1) / 2;
    dim3 grid((unsigned)blocks), blk(SYMV_THREADS);
    switch (b) {
        case 64: symv_partials_kernel<64, true, 4><<<grid, blk, 0, h->stream>>>(A, x, rowpart, colpart, n, nullptr, 0); break;
        case 128: symv_partials_kernel<128, true, 8><<<grid, blk, 0, h->stream>>>(A, x, rowpart, colpart, n, nullptr, 0); break;
        case 256: symv_partials_kernel<256, true, 2><<<grid, blk, 0, h->stream>>>(A, x, rowpart, colpart, n, nullptr, 0); break;
        default: hs_set_error("SYMV: unsupported block size"); return HS_ERR_UNSUPPORTED;
    }
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// blocks = the first `blocks` stored blocks in packed order (0: all of them)
static int launch_partials(hs_handle_t h, const double* A, const double* x, double* rowpart, double* colpart, int n,
                           int b, const i64* stop, i64 it, i64 blocks = 0) {
    if (blocks <= 0) blocks = (i64)n * (n + 1) / 2;
    if (blocks > 0x7fffffffLL) {
        hs_set_error("SYMV: too many blocks");
        return HS_ERR_UNSUPPORTED;
    }
    dim3 grid((unsigned)blocks), blk(SYMV_THREADS);
    switch (b) {
        case 64: symv_partials_kernel<64, false, 4><<<grid, blk, 0, h->stream>>>(A, x, rowpart, colpart, n, stop, it); break;
        case 128: symv_partials_kernel<128, false, 8><<<grid, blk, 0, h->stream>>>(A, x, rowpart, colpart, n, stop, it); break;
        case 256: symv_partials_kernel<256, false, 2><<<grid, blk, 0, h->stream>>>(A, x, rowpart, colpart, n, stop, it); break;
        default: hs_set_error("SYMV: unsupported block size"); return HS_ERR_UNSUPPORTED;
    }
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// workspace layout for SYMV / CG (doubles)
struct SymvWork {
    double* rowpart;
    double* colpart;
    double* part1;
    double* part2;
    double* vec;   // 3 vectors: d, q, r for CG
    CGState* st;
};

static size_t symv_work_bytes(i64 N, int b, bool withVectors) {
    const i64 n = (N + b - 1) / b;
    const i64 nb = n * (n + 1) / 2;
    const i64 rows = n * b;
    const i64 G = rows / FIN_ROWS + 1;  // the largest partial count (reduction pass grid)
    return sizeof(double) * (size_t)(2 * nb * b + 2 * (G + 64) + (withVectors ? 3 * rows : 0)) + sizeof(CGState) +
           256;
}

size_t hs_ws_symv(i64 N, int b) { return symv_fast_ok(b) ? symv_work_bytes(N, b, false) : 0; }
size_t hs_ws_cg(i64 N, int b) { return symv_fast_ok(b) ? symv_work_bytes(N, b, true) : 0; }

// carves the SYMV / CG buffers out of `base` (symv_work_bytes of them)
static void symv_carve(void* base, int n, int b, bool withVectors, SymvWork* w) {
    const i64 nb = (i64)n * (n + 1) / 2;
    const i64 rows = (i64)n * b;
    const i64 G = rows / FIN_ROWS + 1;
    double* p = (double*)base;
    w->rowpart = p;
    p += nb * b;
    w->colpart = p;
    p += nb * b;
    w->part1 = p;
    p += G + 64;
    w->part2 = p;
    p += G + 64;
    w->vec = p;
    if (withVectors) p += 3 * rows;
    w->st = (CGState*)(((uintptr_t)p + 63) & ~(uintptr_t)63);
}

static int symv_workspace(hs_handle_t h, int n, int b, bool withVectors, SymvWork* w) {
    void* base = nullptr;
    int rc = hs_work(h, symv_work_bytes((i64)n * b, b, withVectors), &base);
    if (rc) return rc;
    symv_carve(base, n, b, withVectors, w);
    return HS_OK;
}

extern "C" int hs_symv_packed_f64(hs_handle_t h, const double* A, const double* x, double* y, int64_t N, int b) {
    if (!h) return -1;
    if (!A) return -2;
    if (!x) return -3;
    if (!y) return -4;
    if (N <= 0) return -5;
    if (b <= 0) return -6;
    const int n = (int)((N + b - 1) / b);
    if (!symv_fast_ok(b)) return hs_symv_block_f64(h, A, x, y, b, 0, 0, n, n, n, 1);
    SymvWork w;
    int rc = symv_workspace(h, n, b, false, &w);
    if (rc) return rc;
    if ((rc = launch_partials(h, A, x, w.rowpart, w.colpart, n, b, nullptr, 0))) return rc;
    const i64 rows = (i64)n * b;
    const i64 G = rows / FIN_ROWS;
    symv_finish_kernel<0><<<dim3((unsigned)G), dim3(SYMV_THREADS), 0, h->stream>>>(w.rowpart, w.colpart, n, b, y,
                                                                                   nullptr, nullptr, nullptr, 0);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// Rows [0, blockRows b) of y = A x: the block rows a GPU owns when the rows are
// split between a CPU and a GPU (CG.cpp:43-46). Every block these rows touch
// lies in block columns 0 .. blockRows-1 (blocks (I, J) with J <= I and, read
// transposed, (J, I) with J > I), which is the packed prefix of the first
// S = blockRows n - blockRows (blockRows - 1) / 2 blocks: the read-once pass
// runs over that prefix only, the row sums over those rows only.
extern "C" int hs_symv_rows_f64(hs_handle_t h, const double* A, const double* x, double* y, int64_t N, int b,
                                int blockRows) {
    if (!h) return -1;
    if (!A) return -2;
    if (!x) return -3;
    if (!y) return -4;
    if (N <= 0) return -5;
    if (b <= 0) return -6;
    const int n = (int)((N + b - 1) / b);
    if (blockRows < 0 || blockRows > n) return -7;
    if (blockRows == 0) return HS_OK;
    if (!symv_fast_ok(b)) return hs_symv_block_f64(h, A, x, y, b, 0, 0, blockRows, n, n, 1);
    SymvWork w;
    int rc = symv_workspace(h, n, b, false, &w);
    if (rc) return rc;
    const i64 prefix = (i64)blockRows * n - (i64)blockRows * (blockRows - 1) / 2;
    if ((rc = launch_partials(h, A, x, w.rowpart, w.colpart, n, b, nullptr, 0, prefix))) return rc;
    const i64 G = (i64)blockRows * b / FIN_ROWS;
    symv_finish_kernel<0><<<dim3((unsigned)G), dim3(SYMV_THREADS), 0, h->stream>>>(w.rowpart, w.colpart, n, b, y,
                                                                                   nullptr, nullptr, nullptr, 0);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// hs_symv_packed_f64 repeated `reps` times with HIP events around each pass on
// the handle's stream: pass_ms[0] = average duration of the read-once block
// pass, pass_ms[1] = of the fixed-order row reduction.
extern "C" int hs_symv_packed_f64_timed(hs_handle_t h, const double* A, const double* x, double* y, int64_t N, int b,
                                        int reps, double* pass_ms) {
    if (!h) return -1;
    if (!A) return -2;
    if (!x) return -3;
    if (!y) return -4;
    if (N <= 0) return -5;
    if (!symv_fast_ok(b)) {
        hs_set_error("hs_symv_packed_f64_timed: block size must be 64, 128 or 256");
        return -6;
    }
    if (reps <= 0) return -7;
    if (!pass_ms) return -8;
    const int n = (int)((N + b - 1) / b);
    SymvWork w;
    int rc = symv_workspace(h, n, b, false, &w);
    if (rc) return rc;
    hipEvent_t ev[3];
    for (int e = 0; e < 3; ++e) HS_HIP(hipEventCreate(&ev[e]));  // timing events of this probe entry
    double t0 = 0.0, t1 = 0.0;
    const i64 G = (i64)n * b / FIN_ROWS;
    for (int r = 0; r < reps && rc == HS_OK; ++r) {
        HS_HIP(hipEventRecord(ev[0], h->stream));
        rc = launch_partials(h, A, x, w.rowpart, w.colpart, n, b, nullptr, 0);
        if (rc) break;
        HS_HIP(hipEventRecord(ev[1], h->stream));
        symv_finish_kernel<0><<<dim3((unsigned)G), dim3(SYMV_THREADS), 0, h->stream>>>(
            w.rowpart, w.colpart, n, b, y, nullptr, nullptr, nullptr, 0);
        HS_LAUNCH_CHECK();
        HS_HIP(hipEventRecord(ev[2], h->stream));
        HS_HIP(hipEventSynchronize(ev[2]));
        float a = 0.f, c = 0.f;
        HS_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        HS_HIP(hipEventElapsedTime(&c, ev[1], ev[2]));
        t0 += a;
        t1 += c;
    }
    for (int e = 0; e < 3; ++e) hipEventDestroy(ev[e]);
    if (rc) return rc;
    pass_ms[0] = t0 / reps;
    pass_ms[1] = t1 / reps;
    return HS_OK;
}

// CG at share 1.0 (CG.cpp:38-163, 277-528). The host enqueues iterations in
// chunks; each iteration's kernels check the device stop index first, so
// iterations enqueued past convergence cost only their launch. The stop
// index is read back once per chunk through pinned memory.
static int cg_args(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b, double eps,
                   int64_t imax, const char* who) {
    if (!h) return -1;
    if (!A) return -2;
    if (!rhs) return -3;
    if (!x) return -4;
    if (N <= 0) return -5;
    if (!symv_fast_ok(b)) {
        hs_set_error(std::string(who) + ": block size must be 64, 128 or 256");
        return -6;
    }
    if (!(eps >= 0.0)) return -7;
    if (imax < 0) return -8;
    return HS_OK;
}

// The device-resident CG on buffers carved out of `w` (symv_carve with
// vectors): no allocation, the handle's pre-created events and pinned word.
static int cg_run(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b, double eps,
                  int64_t imax, const SymvWork& w, int64_t* iterations, double* delta_final, double* delta_zero,
                  double* trace = nullptr) {
    const int n = (int)((N + b - 1) / b);
    const i64 rows = (i64)n * b;
    const int G = (int)((rows + SYMV_THREADS - 1) / SYMV_THREADS);  // vector-kernel grid
    const int GF = (int)(rows / FIN_ROWS);                          // reduction-pass grid
    int rc = HS_OK;
    double* d = w.vec;
    double* q = w.vec + rows;
    double* r = w.vec + 2 * rows;
    CGState* st = w.st;
    const i64* stop = &st->stop;
    const dim3 vg((unsigned)G), vb(SYMV_THREADS), fg((unsigned)GF);

    cg_init_kernel<<<vg, vb, 0, h->stream>>>(rhs, x, r, d, rows, w.part2);
    HS_LAUNCH_CHECK();
    cg_init_finish_kernel<<<1, vb, 0, h->stream>>>(w.part2, G, st, eps * eps);
    HS_LAUNCH_CHECK();

    const int CHUNK = 8;
    volatile i64* pinnedStop = (volatile i64*)h->pinned;
    i64 enq = 0;
    i64 stopHost = LLONG_MAX;
    hipEvent_t evs[2] = {h->events[0], h->events[1]};
    int pending = -1;  // slot of the chunk whose stop copy is in flight
    IterTrace tr;
    if ((rc = tr.start(trace, imax, CHUNK))) return rc;
    i64 chunkFirst[2] = {0, 0}, chunkEnd[2] = {0, 0};  // iterations of the chunk in each slot
    auto enqueue_chunk = [&]() -> int {
        const i64 end = std::min<i64>(enq + CHUNK, imax);
        for (; enq < end; ++enq) {
            const i64 it = enq;
            int e;
            if ((e = tr.mark(h->stream, it, 0))) return e;
            if ((e = launch_partials(h, A, d, w.rowpart, w.colpart, n, b, stop, it))) return e;
            symv_finish_kernel<1><<<fg, vb, 0, h->stream>>>(w.rowpart, w.colpart, n, b, q, d, w.part1, stop, it);
            HS_LAUNCH_CHECK();
            if ((e = tr.mark(h->stream, it, 1))) return e;
            const int updateR = (it % 50) != 0;
            cg_xr_kernel<<<vg, vb, 0, h->stream>>>(st, w.part1, GF, d, q, x, r, rows, w.part2, it, updateR);
            HS_LAUNCH_CHECK();
            if ((e = tr.mark(h->stream, it, 2))) return e;
            if (!updateR) {
                // true residual every 50 iterations (CG.cpp:95-101)
                if ((e = launch_partials(h, A, x, w.rowpart, w.colpart, n, b, stop, it))) return e;
                symv_finish_kernel<2><<<fg, vb, 0, h->stream>>>(w.rowpart, w.colpart, n, b, r, rhs, w.part2, stop,
                                                                it);
                HS_LAUNCH_CHECK();
            }
            if ((e = tr.mark(h->stream, it, 3))) return e;
            // r.r partials come from cg_xr (G of them) or from the residual pass (GF)
            cg_d_kernel<<<vg, vb, 0, h->stream>>>(st, w.part2, updateR ? G : GF, r, d, rows, it);
            HS_LAUNCH_CHECK();
            if ((e = tr.mark(h->stream, it, 4))) return e;
        }
        return HS_OK;
    };
    // the executed iterations of the chunk in slot s, once it has completed
    auto read_slot = [&](int s, i64 stopIt) -> int {
        const i64 last = stopIt == LLONG_MAX ? chunkEnd[s] : std::min<i64>(chunkEnd[s], stopIt + 1);
        return tr.read(chunkFirst[s], last);
    };
    int slot = 0;
    while (true) {
        if (enq < imax) {
            chunkFirst[slot] = enq;
            if ((rc = enqueue_chunk())) break;
            chunkEnd[slot] = enq;
            HS_HIP(hipMemcpyAsync((void*)(pinnedStop + slot), stop, sizeof(i64), hipMemcpyDeviceToHost, h->stream));
            HS_HIP(hipEventRecord(evs[slot], h->stream));
        }
        if (pending >= 0) {
            HS_HIP(hipEventSynchronize(evs[pending]));
            stopHost = pinnedStop[pending];
            if ((rc = read_slot(pending, stopHost))) break;
            if (stopHost != LLONG_MAX) break;
        }
        if (enq >= imax) {
            // drain the last chunk's copy
            if (pending != slot) {
                HS_HIP(hipEventSynchronize(evs[slot]));
                stopHost = pinnedStop[slot];
                if ((rc = read_slot(slot, stopHost))) break;
            }
            break;
        }
        pending = slot;
        slot ^= 1;
    }
    if (rc) return rc;
    // the final state through the handle's pinned page (past the two stop slots)
    CGState* pinnedState = (CGState*)((char*)h->pinned + 64);
    HS_HIP(hipMemcpyAsync(pinnedState, st, sizeof(CGState), hipMemcpyDeviceToHost, h->stream));
    HS_HIP(hipStreamSynchronize(h->stream));
    const CGState hs = *pinnedState;
    const i64 iters = (hs.stop == LLONG_MAX) ? enq : hs.stop + 1;
    if (iterations) *iterations = iters;
    if (delta_final) *delta_final = hs.delta[iters & 1];
    if (delta_zero) *delta_zero = hs.delta0;
    return HS_OK;
}

extern "C" int hs_cg_f64(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b, double eps,
                         int64_t imax, int64_t* iterations, double* delta_final, double* delta_zero) {
    int rc = cg_args(h, A, rhs, x, N, b, eps, imax, "hs_cg_f64");
    if (rc) return rc;
    SymvWork w;
    if ((rc = symv_workspace(h, (int)((N + b - 1) / b), b, true, &w))) return rc;
    return cg_run(h, A, rhs, x, N, b, eps, imax, w, iterations, delta_final, delta_zero);
}

extern "C" int hs_cg_f64_trace(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b,
                               double eps, int64_t imax, int64_t* iterations, double* delta_final, double* delta_zero,
                               double* trace) {
    int rc = cg_args(h, A, rhs, x, N, b, eps, imax, "hs_cg_f64_trace");
    if (rc) return rc;
    if (!trace) return -12;
    SymvWork w;
    if ((rc = symv_workspace(h, (int)((N + b - 1) / b), b, true, &w))) return rc;
    return cg_run(h, A, rhs, x, N, b, eps, imax, w, iterations, delta_final, delta_zero, trace);
}

// ---------------------------------------------------------------------------
// Allocation-free CG (hs_cg_f64_noalloc) for callers that cannot allocate
// device memory (graph capture, serving loops). Its scratch, the SYMV
// partials, the three CG vectors and the loop state, lives in `scratch`, a
// file-scope device array of HS_NOALLOC_SCRATCH_BYTES that every GPU the
// library runs on holds until the process ends. On the host the symbol names
// a stand-in, not device memory: hs_noalloc_setup looks up its device address
// once per device (from hs_handle_create, with that device current) and only
// that address is ever handed to kernels. Every call on one GPU shares the
// array: calls that overlap on one GPU race. Nothing in it has to start at
// zero: each call writes every partial and vector it later reads, and its
// init kernels set the loop state.
constexpr size_t HS_NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
constexpr int HS_NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[HS_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::atomic<void*> g_scratch_dev[HS_NOALLOC_MAX_DEVICES];

int hs_noalloc_setup(int device) {
    if (device < 0 || device >= HS_NOALLOC_MAX_DEVICES) {
        hs_set_error("hs_noalloc_setup: device index beyond the supported range");
        return HS_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device].load()) return HS_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hs_hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_dev[device].store(p);
    return HS_OK;
}

extern "C" int hs_cg_noalloc_limit(size_t* scratch_bytes) {
    if (!scratch_bytes) return -1;
    *scratch_bytes = HS_NOALLOC_SCRATCH_BYTES;
    return HS_OK;
}

extern "C" int hs_cg_f64_noalloc(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b,
                                 double eps, int64_t imax, int64_t* iterations, double* delta_final,
                                 double* delta_zero) {
    int rc = cg_args(h, A, rhs, x, N, b, eps, imax, "hs_cg_f64_noalloc");
    if (rc) return rc;
    const size_t need = symv_work_bytes(N, b, true);
    if (need > HS_NOALLOC_SCRATCH_BYTES) {
        hs_set_error("hs_cg_f64_noalloc: N = " + std::to_string(N) + ", b = " + std::to_string(b) + " needs " +
                     std::to_string(need) + " bytes of scratch, beyond its fixed " +
                     std::to_string(HS_NOALLOC_SCRATCH_BYTES) + " (use hs_cg_f64 with a reserved workspace)");
        return HS_ERR_UNSUPPORTED;
    }
    void* base = (h->device >= 0 && h->device < HS_NOALLOC_MAX_DEVICES) ? g_scratch_dev[h->device].load() : nullptr;
    if (!base) {
        hs_set_error("hs_cg_f64_noalloc: no scratch address for this device (handle not created by hs_handle_create)");
        return HS_ERR_UNSUPPORTED;
    }
    SymvWork w;
    symv_carve(base, (int)((N + b - 1) / b), b, true, &w);
    return cg_run(h, A, rhs, x, N, b, eps, imax, w, iterations, delta_final, delta_zero);
}

// Internal SYMV for other solvers: mode 0 y = A x, mode 2 y = rhs - A x (part
// receives rows/64 per-work-group partial sums of y_i^2); abs_val uses |A|.
// rowpart / colpart hold n(n+1)/2 * b doubles each.
int hs_symv_into(hs_handle_t h, const double* A, const double* x, double* y, const double* rhs, i64 N, int b,
                 double* rowpart, double* colpart, double* part, int mode, bool abs_val) {
    if (!symv_fast_ok(b)) {
        hs_set_error("hs_symv_into: block size must be 64, 128 or 256");
        return HS_ERR_UNSUPPORTED;
    }
    const int n = (int)((N + b - 1) / b);
    int rc = abs_val ? launch_partials_abs(h, A, x, rowpart, colpart, n, b)
                     : launch_partials(h, A, x, rowpart, colpart, n, b, nullptr, 0);
    if (rc) return rc;
    const dim3 fg((unsigned)((i64)n * b / FIN_ROWS)), vb(SYMV_THREADS);
    if (mode == 2)
        symv_finish_kernel<2><<<fg, vb, 0, h->stream>>>(rowpart, colpart, n, b, y, rhs, part, nullptr, 0);
    else
        symv_finish_kernel<0><<<fg, vb, 0, h->stream>>>(rowpart, colpart, n, b, y, nullptr, nullptr, nullptr, 0);
    HS_LAUNCH_CHECK();
    return HS_OK;
}
