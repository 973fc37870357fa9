// Blocked Cholesky on gfx950: FP64 MFMA trailing update (SYRK + GEMM fused
// over the packed trailing triangle), the stream-ordered right-looking driver
// and the blocked triangular solves.
//
// Trailing update of step k (Cholesky.cpp:210-259 at share 1.0):
//   A_IJ <- A_IJ - L_Ik L_Jk^T   for k < J <= I < n
// One launch covers every (I,J) pair of the requested rows. A work-group owns
// a TILE x TILE tile of one output block; 4 waves in a 2x2 arrangement each
// own (TILE/2)^2 outputs as 16x16 v_mfma_f64_16x16x4f64 accumulators.
// Operands are staged through LDS k-major ([k][row]) so a wave's 16-lane
// groups read 128 contiguous bytes, in KT = 16 deep stages with the next
// stage's global loads in flight during the MFMAs of the current one.
#include <algorithm>
#include <vector>

#include "hs_internal.h"

typedef double d4 __attribute__((ext_vector_type(4)));

int hs_launch_trailing_generic(hs_handle_t h, double* A, int b, int k, int rowStart, int rowCount, int n, int mode);

namespace {

constexpr int KT = 16;
constexpr int THREADS = 256;

struct PairMap {
    int k, rowStart, mode;
    i64 pairOffset;
};

// decode work-group pair index -> (I, J)
__device__ inline void pair_of(const PairMap& pm, i64 q, int* I, int* J) {
    if (pm.mode == 1) {
        *I = pm.rowStart + (int)q;
        *J = *I;
        return;
    }
    int r, c;
    hs_tri_decode(q + pm.pairOffset, &r, &c);
    if (pm.mode == 2) {
        *I = pm.k + 2 + r;
        *J = pm.k + 1 + c;
    } else {
        *I = pm.k + 1 + r;
        *J = pm.k + 1 + c;
    }
}

template <int TILE>
__global__ __launch_bounds__(THREADS) void trailing_mfma_f64(double* __restrict__ A, int b, int n, PairMap pm) {
    constexpr int WT = TILE / 2;   // wave tile edge
    constexpr int MT = WT / 16;    // 16x16 MFMA tiles per wave edge
    constexpr int LDP = TILE + 4;  // LDS pitch (doubles) of one k row
    __shared__ double As[KT * LDP];
    __shared__ double Bs[KT * LDP];

    const int tilesPerEdge = b / TILE;
    const int tilesPerBlock = tilesPerEdge * tilesPerEdge;
    const i64 q = blockIdx.x / tilesPerBlock;
    const int tile = blockIdx.x % tilesPerBlock;
    const int tm = tile / tilesPerEdge, tn = tile % tilesPerEdge;
    int I, J;
    pair_of(pm, q, &I, &J);
    const bool diag = (I == J);
    if (diag && tm < tn) return;  // strictly upper tile of a diagonal block

    const i64 bb = (i64)b * b;
    const double* La = A + hs_block_id(I, pm.k, n) * bb + (i64)tm * TILE * b;
    const double* Lb = A + hs_block_id(J, pm.k, n) * bb + (i64)tn * TILE * b;
    double* C = A + hs_block_id(I, J, n) * bb + (i64)tm * TILE * b + (i64)tn * TILE;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm0 = (wave >> 1) * WT, wn0 = (wave & 1) * WT;

    // global -> register staging: each thread loads LPT doubles of one row
    constexpr int LPT = TILE * KT / THREADS;  // 8 (TILE 128) or 4 (TILE 64)
    constexpr int TPR = KT / LPT;             // threads per row
    const int lrow = tid / TPR;
    const int lk = (tid % TPR) * LPT;
    double ra[LPT], rb[LPT];

    d4 acc[MT][MT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    const int nk = b / KT;
    auto gload = [&](int kt) {
        const double* pa = La + (i64)lrow * b + kt * KT + lk;
        const double* pb = Lb + (i64)lrow * b + kt * KT + lk;
#pragma unroll
        for (int v = 0; v < LPT; v += 2) {
            const double2 x = *reinterpret_cast<const double2*>(pa + v);
            const double2 y = *reinterpret_cast<const double2*>(pb + v);
            ra[v] = x.x;
            ra[v + 1] = x.y;
            rb[v] = y.x;
            rb[v + 1] = y.y;
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int v = 0; v < LPT; ++v) {
            As[(lk + v) * LDP + lrow] = ra[v];
            Bs[(lk + v) * LDP + lrow] = rb[v];
        }
    };

    gload(0);
    sstore();
    __syncthreads();
    const int lr = lane & 15, lq = lane >> 4;
    for (int kt = 0; kt < nk; ++kt) {
        if (kt + 1 < nk) gload(kt + 1);
#pragma unroll
        for (int ks = 0; ks < KT; ks += 4) {
            double av[MT], bv[MT];
#pragma unroll
            for (int i = 0; i < MT; ++i) av[i] = As[(ks + lq) * LDP + wm0 + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < MT; ++j) bv[j] = Bs[(ks + lq) * LDP + wn0 + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < MT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
        if (kt + 1 < nk) {
            sstore();
            __syncthreads();
        }
    }
    // epilogue: C -= acc. For the f64 16x16x4 MFMA, register v of lane l holds
    // row l/16 + 4v, column l%16 of the 16x16 tile (one row group per register,
    // unlike the f32 variant's four consecutive rows).
#pragma unroll
    for (int i = 0; i < MT; ++i) {
#pragma unroll
        for (int j = 0; j < MT; ++j) {
            const int col = wn0 + j * 16 + lr;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wm0 + i * 16 + lq + 4 * v;
                if (diag && tm == tn && row < col) continue;
                double* p = C + (i64)row * b + col;
                *p = *p - acc[i][j][v];
            }
        }
    }
}

}  // namespace

// Dispatch: MFMA kernel when b is a multiple of 64 (the reference's optimised
// GEMM has the same constraint, MatrixMatrixOperations.cpp:657-659), else the
// element-wise kernel.
int hs_launch_trailing(hs_handle_t h, double* A, int b, int k, int rowStart, int rowCount, int n, int mode) {
    if (rowCount <= 0) return HS_OK;
    if (b % 64 != 0) return hs_launch_trailing_generic(h, A, b, k, rowStart, rowCount, n, mode);
    const int tile = (b % 128 == 0) ? 128 : 64;
    const i64 tilesPerBlock = (i64)(b / tile) * (b / tile);
    auto launch = [&](int md, i64 pairs, i64 off, int rs) -> int {
        if (pairs <= 0) return HS_OK;
        PairMap pm{k, rs, md, off};
        const i64 grid = pairs * tilesPerBlock;
        if (grid > 0x7fffffffLL) {
            hs_set_error("trailing update grid too large");
            return HS_ERR_UNSUPPORTED;
        }
        if (tile == 128)
            trailing_mfma_f64<128><<<dim3((unsigned)grid), dim3(THREADS), 0, h->stream>>>(A, b, n, pm);
        else
            trailing_mfma_f64<64><<<dim3((unsigned)grid), dim3(THREADS), 0, h->stream>>>(A, b, n, pm);
        HS_LAUNCH_CHECK();
        return HS_OK;
    };
    if (mode == 1) return launch(1, rowCount, 0, rowStart);
    if (mode == 2) {
        const int firstRow = std::max(rowStart, k + 2);
        const int lastRow = rowStart + rowCount;
        if (lastRow <= firstRow) return HS_OK;
        const i64 r0 = firstRow - (k + 2), r1 = lastRow - (k + 2);
        return launch(2, r1 * (r1 + 1) / 2 - r0 * (r0 + 1) / 2, r0 * (r0 + 1) / 2, rowStart);
    }
    // mode 3: every (I,J) with I in the row range and k < J <= I
    const i64 r0 = rowStart - (k + 1), r1 = rowStart + rowCount - (k + 1);
    return launch(3, r1 * (r1 + 1) / 2 - r0 * (r0 + 1) / 2, r0 * (r0 + 1) / 2, rowStart);
}

// ---------------------------------------------------------------------------
// Whole factorisation, Cholesky::solve_heterogeneous at share 1.0
// (Cholesky.cpp:446-476): per column k, POTRF of the diagonal block, TRSM of
// the n-k-1 blocks below it, then the fused SYRK+GEMM trailing update.
// Everything is stream ordered; the only host synchronisation is the final
// read of the pivot status.
// ---------------------------------------------------------------------------
extern "C" int hs_cholesky_f64(hs_handle_t h, double* A, int64_t N, int b) {
    if (!h) return -1;
    if (!A) return -2;
    if (N <= 0) return -3;
    if (b <= 0) return -4;
    const int n = (int)((N + b - 1) / b);
    int rc = hs_reset_info(h);
    if (rc) return rc;
    for (int k = 0; k < n; ++k) {
        const i64 id = hs_block_id(k, k, n);
        if ((rc = hs_launch_potrf(h, A, N, b, id, k))) return rc;
        const int cnt = n - k - 1;
        if (cnt > 0) {
            if ((rc = hs_launch_trsm(h, A, b, id, k, k + 1, cnt))) return rc;
            if ((rc = hs_launch_trailing(h, A, b, k, k + 1, cnt, n, 3))) return rc;
        }
    }
    int64_t info = 0;
    if ((rc = hs_get_info(h, &info))) return rc;
    return (int)info;
}

extern "C" int hs_cholesky_f64_timed(hs_handle_t h, double* A, int64_t N, int b, double* phase_ms,
                                     int64_t* launches) {
    if (!h) return -1;
    if (!A) return -2;
    if (N <= 0) return -3;
    if (b <= 0) return -4;
    if (!phase_ms) return -5;
    if (!launches) return -6;
    const int n = (int)((N + b - 1) / b);
    int rc = hs_reset_info(h);
    if (rc) return rc;
    // events: before POTRF, before TRSM, before update, after update, per step
    std::vector<hipEvent_t> ev((size_t)4 * n);
    for (auto& e : ev) HS_HIP(hipEventCreate(&e));
    for (int k = 0; k < n; ++k) {
        const i64 id = hs_block_id(k, k, n);
        const int cnt = n - k - 1;
        HS_HIP(hipEventRecord(ev[4 * k], h->stream));
        if ((rc = hs_launch_potrf(h, A, N, b, id, k))) break;
        HS_HIP(hipEventRecord(ev[4 * k + 1], h->stream));
        if (cnt > 0 && (rc = hs_launch_trsm(h, A, b, id, k, k + 1, cnt))) break;
        HS_HIP(hipEventRecord(ev[4 * k + 2], h->stream));
        if (cnt > 0 && (rc = hs_launch_trailing(h, A, b, k, k + 1, cnt, n, 3))) break;
        HS_HIP(hipEventRecord(ev[4 * k + 3], h->stream));
    }
    if (rc == 0) {
        HS_HIP(hipStreamSynchronize(h->stream));
        double t[3] = {0, 0, 0};
        for (int k = 0; k < n; ++k)
            for (int p = 0; p < 3; ++p) {
                float ms = 0;
                HS_HIP(hipEventElapsedTime(&ms, ev[4 * k + p], ev[4 * k + p + 1]));
                t[p] += ms;
            }
        for (int p = 0; p < 3; ++p) phase_ms[p] = t[p];
        launches[0] = n;
        launches[1] = n - 1;
        launches[2] = n - 1;
    }
    for (auto& e : ev) hipEventDestroy(e);
    if (rc) return rc;
    int64_t info = 0;
    if ((rc = hs_get_info(h, &info))) return rc;
    return (int)info;
}

// Forward and backward blocked substitution (TriangularSystemSolver.cpp:31-82).
extern "C" int hs_tri_solve_f64(hs_handle_t h, const double* L, double* x, int64_t N, int b) {
    if (!h) return -1;
    if (!L) return -2;
    if (!x) return -3;
    if (N <= 0) return -4;
    if (b <= 0) return -5;
    const int n = (int)((N + b - 1) / b);
    int rc;
    for (int j = 0; j < n; ++j) {
        const i64 id = hs_block_id(j, j, n);
        if ((rc = hs_launch_trsv_block(h, L, x, N, b, j, id, 0))) return rc;
        if (j < n - 1 && (rc = hs_launch_col_update(h, L, x, b, j + 1, n - (j + 1), j, id, n, 0))) return rc;
    }
    for (int j = n - 1; j >= 0; --j) {
        const i64 id = hs_block_id(j, j, n);
        if ((rc = hs_launch_trsv_block(h, L, x, N, b, j, id, 1))) return rc;
        if (j > 0 && (rc = hs_launch_col_update(h, L, x, b, 0, j, j, id, n, 1))) return rc;
    }
    return HS_OK;
}
