// Panel kernels of the blocked Cholesky for the solver block sizes
// (b % 32 == 0 for POTRF, b % 16 == 0 for TRSM, b <= 256), FP64 arithmetic
// on FP64 or FP32 storage. Both sit on the factorisation's critical path, so
// they are built for latency and for a small resource footprint (they run on
// CUs shared with the trailing-update work-groups):
//
//  potrf_rows_kernel   one work-group, left-looking over 32-wide panels: the
//                      update of a panel by the finished columns runs on
//                      v_mfma_f64_16x16x4f64 with operands loaded straight from
//                      the block (32 B per lane), the panel is then factored
//                      with thread t holding row t in registers, one FMA per
//                      entry and column and one barrier per column. It also
//                      writes the inverses of the 16x16 diagonal sub-blocks
//                      of the factor for the TRSM.
//  trsm_inv_kernel     B <- B L^{-T} for every panel block: one wave owns 16
//                      rows of one block and solves L X^T = B^T by 16-row block
//                      substitution entirely in registers. The MFMA output
//                      layout of one step is the B-operand layout of the next
//                      (with the row permutation below), so nothing goes
//                      through LDS.
//  tri_inv16_kernel    the 16x16 diagonal sub-block inverses from a stored
//                      factor, for a TRSM that does not follow this POTRF.
//
// Reference: MatrixOperations::cholesky_optimizedGPU (MatrixOperations.cpp:
// 105-190) and MatrixMatrixOperations::triangularSolve_optimizedGPU
// (MatrixMatrixOperations.cpp:50-107).
#include <algorithm>

#include "hs_internal.h"

typedef double pd4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int PW = 32;       // POTRF panel width
constexpr int PP = PW + 1;   // LDS pitch of a panel row

// 1 / sqrt(d) for d > 0 to within an ulp or two: v_rsq_f64 (relative error
// below 2^-22) and two Newton steps y += y (1 - d y^2) / 2, each squaring the
// error. Shorter on the dependent path than an IEEE square root followed by
// an IEEE division.
__device__ inline double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double r = fma(-(d * y), y, 1.0);
        y = fma(0.5 * y, r, y);
    }
    return y;
}

// four consecutive stored values as doubles
__device__ inline void load4(const double* p, double* o) {
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 c = *reinterpret_cast<const double2*>(p + 2);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = c.x;
    o[3] = c.y;
}
__device__ inline void load4(const float* p, double* o) {
    const float4 f = *reinterpret_cast<const float4*>(p);
    o[0] = f.x;
    o[1] = f.y;
    o[2] = f.z;
    o[3] = f.w;
}
__device__ inline void store4(double* p, const double* v) {
    *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
}
__device__ inline void store4(float* p, const double* v) {
    *reinterpret_cast<float4*>(p) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}

// Row permutation that turns the f64 MFMA output layout into its own B-operand
// layout: a 16x16 result tile keeps row g + 4v of lane group g in register v;
// feeding the A operand with its rows permuted by perm16 makes register v of
// group g hold logical row 4g + v, i.e. four consecutive rows per lane, which
// is what one MFMA step over logical k = {4g + v} reads as its B operand.
__device__ inline int perm16(int i) { return 4 * (i & 3) + (i >> 2); }

template <typename ST>
__global__ __launch_bounds__(256) void potrf_rows_kernel(ST* __restrict__ A, i64 N, int b, i64 blockID, int blockRow,
                                                         int* __restrict__ info, double* __restrict__ W) {
    extern __shared__ double smem[];
    double* P = smem;                   // [b][PP] panel
    double* bc = smem + (size_t)b * PP; // [2][PW] column broadcast
    ST* B = A + blockID * (i64)b * b;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const i64 rowBase = (i64)blockRow * b;
    const int nvalid = (int)std::min<i64>((i64)b, std::max<i64>((i64)0, N - rowBase));

    for (int p = 0; p < b; p += PW) {
        const int rows = b - p;
        const int rtiles = rows >> 4;
        // (1) left-looking update acc = L[p+r, 0:p] L[p+c, 0:p]^T on MFMA; wave w
        // owns row tiles w, w+4, w+8, w+12 and both 16-column tiles. Per 16-deep
        // k group a lane holds 4 consecutive k (4 lq .. 4 lq + 3); MFMA step s
        // contracts k = 4 lq + s over the lane groups, the same for A and B.
        pd4 acc[4][2];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i][0] = acc[i][1] = pd4{0.0, 0.0, 0.0, 0.0};
        if (p > 0) {
            double ca[4][4], cb[2][4], na[4][4], nb[2][4];
            auto gload = [&](int k0, double (&a)[4][4], double (&bb)[2][4]) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int rt = wave + 4 * i;
                    if (rt < rtiles) load4(B + (i64)(p + rt * 16 + lr) * b + k0 + 4 * lq, a[i]);
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) load4(B + (i64)(p + j * 16 + lr) * b + k0 + 4 * lq, bb[j]);
            };
            gload(0, ca, cb);
            for (int k0 = 0; k0 < p; k0 += 16) {
                if (k0 + 16 < p) gload(k0 + 16, na, nb);
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if (wave + 4 * i < rtiles) {
                            acc[i][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[i][s], cb[0][s], acc[i][0], 0, 0, 0);
                            acc[i][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[i][s], cb[1][s], acc[i][1], 0, 0, 0);
                        }
                    }
                if (k0 + 16 < p) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
#pragma unroll
                        for (int s = 0; s < 4; ++s) ca[i][s] = na[i][s];
#pragma unroll
                    for (int j = 0; j < 2; ++j)
#pragma unroll
                        for (int s = 0; s < 4; ++s) cb[j][s] = nb[j][s];
                }
            }
        }
        // (2) panel = stored values - acc, lower part; result register v of lane
        // l is row l/16 + 4v, column l%16 of its tile. Padding rows (>= N) keep
        // their stored values.
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int rt = wave + 4 * i;
            if (rt < rtiles) {
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = rt * 16 + lq + 4 * v, c = j * 16 + lr;
                        double val = 0.0;
                        if (r >= c) {
                            val = (double)B[(i64)(p + r) * b + p + c];
                            if (p + r < nvalid) val -= acc[i][j][v];
                        }
                        P[r * PP + c] = val;
                    }
            }
        }
        __syncthreads();
        // (3) unblocked factorisation of the tall panel, thread t = row t
        const int t = tid;
        const bool have = t < rows;
        const bool valid = (p + t) < nvalid;
        double x[PW];
#pragma unroll
        for (int c = 0; c < PW; ++c) x[c] = have ? P[t * PP + c] : 0.0;
        // One column per barrier. Rows c .. PW-1 publish their column-c
        // value; with s = sqrt(A_cc), the row's multiplier a = L_tc / s
        // turns L_tc L_jc = a A_jc into one FMA on the published value, and
        // 1 / s comes from rsqrt_nr. A padding pivot (p + c >= nvalid) is
        // the square root of the stored diagonal; padding rows keep their
        // stored values.
#pragma unroll
        for (int c = 0; c < PW; ++c) {
            double* bcb = bc + (c & 1) * PW;
            if (t >= c && t < PW) bcb[t] = x[c];
            __syncthreads();
            const double d = bcb[c];
            const bool vc = p + c < nvalid;
            if (t == 0 && vc && !(d > 0.0)) atomicCAS(info, 0, (int)(rowBase + p + c + 1));
            if (have) {
                if (t == c) {
                    x[c] = sqrt(d);
                } else if (t > c && valid) {
                    const double inv = rsqrt_nr(d);
                    const double l = x[c] * inv;  // L_tc
                    const double a = l * inv;
                    x[c] = l;
#pragma unroll
                    for (int cc = c + 1; cc < PW; ++cc) x[cc] = fma(-a, bcb[cc], x[cc]);
                }
            }
        }
        // (4) write back; zero the upper triangle of valid rows
        if (have) {
            const i64 gi = p + t;
            if (valid) {
                ST* dst = B + gi * b + p;
#pragma unroll
                for (int c = 0; c < PW; c += 4) {
                    double v4[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) v4[e] = (t >= c + e) ? x[c + e] : 0.0;
                    store4(dst + c, v4);
                }
            } else if (t < PW) {
#pragma unroll
                for (int c = 0; c < PW; ++c)
                    if (t == c) B[gi * b + p + c] = (ST)x[c];
            }
        }
        for (int r = 0; r < PW; ++r) {
            if (p + r >= nvalid) break;
            for (int j = p + PW + tid; j < b; j += 256) B[(i64)(p + r) * b + j] = (ST)0;
        }
        __syncthreads();
    }
    // (5) inverses of the 16x16 diagonal sub-blocks of the stored factor,
    // W[tile][i][j], one column j of tile tb per thread 16 tb + j, all tiles
    // at once (the stores of every panel are visible to the work-group after
    // the barrier above); zero past N. The reciprocals of the diagonal come
    // first, all independent, so that no division sits on the substitution's
    // dependent path.
    if (W && tid < 16 * (b / 16)) {
        const int tb = tid >> 4, j = tid & 15, base = 16 * tb;
        const ST* Ls = B + (i64)base * b + base;
        double w[16], rd[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) rd[i] = 1.0 / (double)Ls[(i64)i * b + i];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            double a = 0.0;
#pragma unroll
            for (int k = 0; k < i; ++k) a += (double)Ls[(i64)i * b + k] * w[k];
            const double num = (i == j) ? 1.0 : -a;
            w[i] = (i < j) ? 0.0 : num * rd[i];
        }
        double* Wt = W + (i64)tb * 256;
        const bool colValid = base + j < nvalid;
#pragma unroll
        for (int i = 0; i < 16; ++i) Wt[i * 16 + j] = (colValid && base + i < nvalid) ? w[i] : 0.0;
    }
}

// Inverses of the 16x16 diagonal sub-blocks of a stored factor block; work-
// group tb handles sub-block tb, thread j < 16 its column j.
template <typename ST>
__global__ __launch_bounds__(64) void tri_inv16_kernel(const ST* __restrict__ A, int b, i64 blockID,
                                                       double* __restrict__ W) {
    const int tb = blockIdx.x;
    const int j = threadIdx.x;
    if (j >= 16) return;
    const ST* L = A + blockID * (i64)b * b + (i64)(16 * tb) * b + 16 * tb;
    double w[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k < i; ++k) a += (double)L[(i64)i * b + k] * w[k];
        const double num = (i == j) ? 1.0 : -a;
        w[i] = (i < j) ? 0.0 : num / (double)L[(i64)i * b + i];
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) W[(i64)tb * 256 + i * 16 + j] = w[i];
}

// X L^T = B for the panel blocks I = blockStart .. blockStart + count - 1 of
// column blockRow; work-group g (one wave) owns rows 16 (g % NT) .. + 15 of
// block blockStart + g / NT. Lane l: X row c = l % 16, group q = l / 16.
// X^T tile t (16 logical rows) sits in 4 registers, register v of group q
// holding logical row 4q + v (see perm16).
template <int NT, typename ST>
__global__ __launch_bounds__(64) void trsm_inv_kernel(ST* __restrict__ A, i64 blockID, int blockRow, int blockStart,
                                                      const double* __restrict__ W) {
    constexpr int b = 16 * NT;
    const int blk = blockIdx.x / NT;
    const int r0 = (blockIdx.x % NT) * 16;
    const int lane = threadIdx.x;
    const int c = lane & 15, q = lane >> 4;
    const int pr = perm16(c);  // logical A-operand row of this lane
    const ST* L = A + blockID * (i64)b * b;
    ST* Bm = A + (blockID + (blockStart - blockRow) + blk) * (i64)b * b;

    pd4 X[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        double v4[4];
        load4(Bm + (i64)(r0 + c) * b + 16 * t + 4 * q, v4);
        X[t] = pd4{v4[0], v4[1], v4[2], v4[3]};
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        // X_t <- W_t X_t
        double wa[4];
        load4(W + (i64)t * 256 + pr * 16 + 4 * q, wa);
        pd4 y = pd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int v = 0; v < 4; ++v) y = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[v], X[t][v], y, 0, 0, 0);
        X[t] = y;
        // X_u <- X_u - L_ut X_t for the later tiles
#pragma unroll
        for (int u = t + 1; u < NT; ++u) {
            double la[4];
            load4(L + (i64)(16 * u + pr) * b + 16 * t + 4 * q, la);
#pragma unroll
            for (int v = 0; v < 4; ++v) X[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-la[v], X[t][v], X[u], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const double v4[4] = {X[t][0], X[t][1], X[t][2], X[t][3]};
        store4(Bm + (i64)(r0 + c) * b + 16 * t + 4 * q, v4);
    }
}

template <int NT, typename ST>
void launch_trsm_nt(hipStream_t s, ST* A, i64 blockID, int blockRow, int blockStart, int blockCount, const double* W) {
    trsm_inv_kernel<NT, ST><<<dim3((unsigned)(blockCount * NT)), dim3(64), 0, s>>>(A, blockID, blockRow, blockStart, W);
}

}  // namespace

bool hs_potrf_fast_ok(int b) { return b % 32 == 0 && b >= 32 && b <= 256; }
bool hs_trsm_fast_ok(int b) { return b % 16 == 0 && b >= 16 && b <= 256; }

template <typename ST>
int hs_launch_potrf_fast(hs_handle_t h, ST* A, i64 N, int b, i64 blockID, int blockRow, double* W, size_t minLds) {
    if (!hs_potrf_fast_ok(b)) {
        hs_set_error("panel POTRF needs b % 32 == 0 and b <= 256");
        return HS_ERR_UNSUPPORTED;
    }
    const size_t lds = std::max(((size_t)b * PP + 2 * PW) * sizeof(double), minLds);
    if (lds > 64 * 1024)
        HS_HIP(hipFuncSetAttribute((const void*)potrf_rows_kernel<ST>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    potrf_rows_kernel<ST><<<1, 256, lds, h->stream>>>(A, N, b, blockID, blockRow, h->info_dev, W);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

template <typename ST>
int hs_launch_trsm_fast(hs_handle_t h, ST* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount,
                        const double* W) {
    if (blockCount <= 0) return HS_OK;
    if (!hs_trsm_fast_ok(b)) {
        hs_set_error("panel TRSM needs b % 16 == 0 and b <= 256");
        return HS_ERR_UNSUPPORTED;
    }
    if (!W) {
        // no POTRF of this block left its sub-block inverses behind
        tri_inv16_kernel<ST><<<dim3((unsigned)(b / 16)), dim3(64), 0, h->stream>>>(A, b, blockID, h->trsm_w);
        HS_LAUNCH_CHECK();
        W = h->trsm_w;
    }
    hipStream_t s = h->stream;
    switch (b / 16) {
        case 1: launch_trsm_nt<1>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 2: launch_trsm_nt<2>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 3: launch_trsm_nt<3>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 4: launch_trsm_nt<4>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 5: launch_trsm_nt<5>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 6: launch_trsm_nt<6>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 7: launch_trsm_nt<7>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 8: launch_trsm_nt<8>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 9: launch_trsm_nt<9>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 10: launch_trsm_nt<10>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 11: launch_trsm_nt<11>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 12: launch_trsm_nt<12>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 13: launch_trsm_nt<13>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 14: launch_trsm_nt<14>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        case 15: launch_trsm_nt<15>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
        default: launch_trsm_nt<16>(s, A, blockID, blockRow, blockStart, blockCount, W); break;
    }
    HS_LAUNCH_CHECK();
    return HS_OK;
}

template <>
int hs_factor_potrf<double>(hs_handle_t h, double* A, i64 N, int b, i64 blockID, int blockRow, size_t minLds) {
    if (hs_potrf_fast_ok(b)) return hs_launch_potrf_fast<double>(h, A, N, b, blockID, blockRow, h->trsm_w, minLds);
    return hs_launch_potrf(h, A, N, b, blockID, blockRow);
}
template <>
int hs_factor_potrf<float>(hs_handle_t h, float* A, i64 N, int b, i64 blockID, int blockRow, size_t minLds) {
    return hs_launch_potrf_fast<float>(h, A, N, b, blockID, blockRow, h->trsm_w, minLds);
}
template <>
int hs_factor_trsm<double>(hs_handle_t h, double* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount) {
    if (hs_trsm_fast_ok(b))
        return hs_launch_trsm_fast<double>(h, A, b, blockID, blockRow, blockStart, blockCount,
                                           hs_potrf_fast_ok(b) ? h->trsm_w : nullptr);
    return hs_launch_trsm(h, A, b, blockID, blockRow, blockStart, blockCount);
}
template <>
int hs_factor_trsm<float>(hs_handle_t h, float* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount) {
    return hs_launch_trsm_fast<float>(h, A, b, blockID, blockRow, blockStart, blockCount,
                                      hs_potrf_fast_ok(b) ? h->trsm_w : nullptr);
}

template int hs_launch_potrf_fast<double>(hs_handle_t, double*, i64, int, i64, int, double*, size_t);
template int hs_launch_potrf_fast<float>(hs_handle_t, float*, i64, int, i64, int, double*, size_t);
template int hs_launch_trsm_fast<double>(hs_handle_t, double*, int, i64, int, int, int, const double*);
template int hs_launch_trsm_fast<float>(hs_handle_t, float*, int, i64, int, int, int, const double*);
