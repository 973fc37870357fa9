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
#include <cmath>
#include <cstdlib>
#include <vector>

#include "hs_factor.h"
#include "hs_internal.h"

typedef double d4 __attribute__((ext_vector_type(4)));

int hs_launch_trailing_generic(hs_handle_t h, double* A, int b, int k, int rowStart, int rowCount, int n, int mode);

namespace {

constexpr int KT = 16;
constexpr int THREADS = 256;

// KIND only separates the look-ahead column launches (KIND 1, modes 5 and 6)
// from the rest in profiles; the code is the same.
template <int TILE, int KIND>
__global__ __launch_bounds__(THREADS) void trailing_mfma_f64(double* __restrict__ A, int b, int n, PairMap pm) {
    constexpr int WT = TILE / 2;   // wave tile edge
    constexpr int MT = WT / 16;    // 16x16 MFMA tiles per wave edge
    constexpr int LDP = TILE + 4;  // LDS pitch (doubles) of one k row
    constexpr int STAGE = KT * LDP;
    __shared__ double As[2 * STAGE];
    __shared__ double Bs[2 * STAGE];

    const int tilesPerEdge = b / TILE;
    const int tilesPerBlock = tilesPerEdge * tilesPerEdge;
    int I, J, tile;
    if (!tile_of(pm, tilesPerBlock, &I, &J, &tile)) return;
    const int tm = tile / tilesPerEdge, tn = tile % tilesPerEdge;
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
    double2 ra[LPT / 2], rb[LPT / 2];
    const double* pa = La + (i64)lrow * b + lk;
    const double* pb = Lb + (i64)lrow * b + lk;

    d4 acc[MT][MT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    // np panels: in column-major block order, block (I, k+p) lies
    // p (n-k-1) - p (p-1)/2 blocks after block (I, k), the same for every row
    const int nk = b / KT, nkt = nk * pm.np;
    auto poff = [&](int kt) -> i64 {
        const i64 pp = kt / nk;
        return (pp * (n - pm.k - 1) - pp * (pp - 1) / 2) * bb + (i64)(kt % nk) * KT;
    };
    auto gload = [&](int kt) {
#pragma unroll
        for (int v = 0; v < LPT / 2; ++v) {
            ra[v] = *reinterpret_cast<const double2*>(pa + poff(kt) + 2 * v);
            rb[v] = *reinterpret_cast<const double2*>(pb + poff(kt) + 2 * v);
        }
    };
    auto sstore = [&](int buf) {
        double* as = As + buf * STAGE;
        double* bs = Bs + buf * STAGE;
#pragma unroll
        for (int v = 0; v < LPT / 2; ++v) {
            as[(lk + 2 * v) * LDP + lrow] = ra[v].x;
            as[(lk + 2 * v + 1) * LDP + lrow] = ra[v].y;
            bs[(lk + 2 * v) * LDP + lrow] = rb[v].x;
            bs[(lk + 2 * v + 1) * LDP + lrow] = rb[v].y;
        }
    };

    // pipeline: stage kt computes from LDS buffer kt&1 while stage kt+1 sits
    // in the other buffer and stage kt+2 is in flight from global memory.
    // The other buffer was last read in stage kt-1, which every wave left at
    // the barrier, so stage kt+1 is written to it right after the first
    // k-step's fragment reads: the LDS stores and the next global loads then
    // issue between this stage's MFMAs (3.5% faster than after them).
    gload(0);
    sstore(0);
    if (nkt > 1) gload(1);
    __syncthreads();
    const int lr = lane & 15, lq = lane >> 4;
    for (int kt = 0; kt < nkt; ++kt) {
        const double* as = As + (kt & 1) * STAGE;
        const double* bs = Bs + (kt & 1) * STAGE;
#pragma unroll
        for (int ks = 0; ks < KT; ks += 4) {
            double av[MT], bv[MT];
#pragma unroll
            for (int i = 0; i < MT; ++i) av[i] = as[(ks + lq) * LDP + wm0 + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < MT; ++j) bv[j] = bs[(ks + lq) * LDP + wn0 + j * 16 + lr];
            if (ks == 0 && kt + 1 < nkt) {
                sstore((kt + 1) & 1);
                if (kt + 2 < nkt) gload(kt + 2);
            }
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < MT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // epilogue: C -= acc as fire-and-forget FP64 atomic adds of -acc (one
    // per element, so the result is the plain fl(C - acc), bit for bit):
    // no load of C and no wait for one, so the work-group leaves the CU
    // (to the other resident one) as soon as the adds are issued; about
    // 1.5% faster than reading C in 16-row batches and storing. For the f64
    // 16x16x4 MFMA, register v of lane l holds row l/16 + 4v, column l%16 of
    // the 16x16 tile. On the diagonal tile of a diagonal block only the lower
    // triangle is written (the reference's SYRK writes i >= j only,
    // MatrixMatrixOperations.cpp:273-299). A must be device memory of the
    // ordinary (coarse-grained) kind, where these atomics are performed in
    // the L2.
    const bool dtile = diag && tm == tn;
    double* pc = C + (i64)(wm0 + lq) * b + wn0 + lr;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wm0 + i * 16 + lq + 4 * v, col = wn0 + j * 16 + lr;
                if (!dtile || row >= col) unsafeAtomicAdd(pc + (i64)(i * 16 + 4 * v) * b + j * 16, -acc[i][j][v]);
            }
}

}  // namespace

// A half-CU launch asks for HALF_CU_LDS bytes of dynamic LDS it does not use:
// 67.5 + 16 KB per work-group keeps a second one off the CU (160 KB), while a
// column update (67.5 KB) or TRSM still fits beside it. A POTRF launched
// beside half-CU updates asks for HALF_CU_POTRF_LDS, more than the free half,
// so that it starts only on a CU of its own.
constexpr int HALF_CU_LDS = 16 * 1024;
constexpr size_t HALF_CU_POTRF_LDS = 80 * 1024;

// Dispatch: MFMA kernel when b is a multiple of 64 (the reference's optimised
// GEMM has the same constraint, MatrixMatrixOperations.cpp:657-659), else the
// element-wise kernel.

int hs_launch_trailing_on(hs_handle_t h, hipStream_t stream, double* A, int b, int k, int rowStart, int rowCount, int n,
                          int mode, int np, bool halfCu) {
    if (rowCount <= 0) return HS_OK;
    if (b % 64 != 0) {
        if (mode > 3) {
            hs_set_error("look-ahead trailing modes need b % 64 == 0");
            return HS_ERR_UNSUPPORTED;
        }
        return hs_launch_trailing_generic(h, A, b, k, rowStart, rowCount, n, mode);
    }
    const int tile = (b % 128 == 0) ? 128 : 64;
    const i64 tilesPerBlock = (i64)(b / tile) * (b / tile);
    PairMap pm;
    const i64 grid = hs_pair_map(k, rowStart, rowCount, n, mode, np, tilesPerBlock, &pm);
    if (grid <= 0) return HS_OK;
    if (grid > 0x7fffffffLL) {
        hs_set_error("trailing update grid too large");
        return HS_ERR_UNSUPPORTED;
    }
    const dim3 g((unsigned)grid), t(THREADS);
    if (tile == 128) {
        if (mode >= 5) {
            trailing_mfma_f64<128, 1><<<g, t, 0, stream>>>(A, b, n, pm);
        } else if (halfCu) {
            // set per launch: the attribute belongs to the device current at the call
            HS_HIP(hipFuncSetAttribute((const void*)trailing_mfma_f64<128, 0>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, HALF_CU_LDS));
            trailing_mfma_f64<128, 0><<<g, t, HALF_CU_LDS, stream>>>(A, b, n, pm);
        } else {
            trailing_mfma_f64<128, 0><<<g, t, 0, stream>>>(A, b, n, pm);
        }
    } else {
        if (mode >= 5) trailing_mfma_f64<64, 1><<<g, t, 0, stream>>>(A, b, n, pm);
        else trailing_mfma_f64<64, 0><<<g, t, 0, stream>>>(A, b, n, pm);
    }
    HS_LAUNCH_CHECK();
    return HS_OK;
}

int hs_launch_trailing(hs_handle_t h, double* A, int b, int k, int rowStart, int rowCount, int n, int mode) {
    return hs_launch_trailing_on(h, h->stream, A, b, k, rowStart, rowCount, n, mode, 1);
}

// Rest updates of at most this many block rows run half-CU (one work-group
// per CU). A step's rest update takes P b^3 m^2 flops for m rows; its panel
// chain P (POTRF + TRSM + one-panel column update), about 1.1 us per unit of
// b for each panel when those kernels run on a CU of their own, and up to
// three times that beside full-CU update work-groups. Half-CU costs the
// update a fifth of its rate; measured best (GP NLL at N = 16384, b = 128:
// NLL 39.2 -> 37.6 ms) with the bound b^3 m^2 / 60 TFLOP/s < 2.5 * 1.1 us * b,
// i.e. m <= 100 at b = 128 and m <= 50 at b = 256 (no change at N = 65536).
// HS_HALF_CU_ROWS sets another bound (0 turns it off).
static int hs_half_cu_rows(int b) {
    static const int env = [] {
        const char* e = std::getenv("HS_HALF_CU_ROWS");
        return e ? std::max(0, std::atoi(e)) : -1;
    }();
    if (env >= 0) return env;
    if (b % 128 != 0) return 0;
    const double m2 = 60e12 * 2.5 * 1.1e-6 * b / ((double)b * b * b);
    return (int)std::sqrt(m2);
}

// Look-ahead FP64 factorisation (schedule in hs_factor.h).
static int cholesky_lookahead(hs_handle_t h, double* A, i64 N, int b, PhaseTimer& pt) {
    const int n = (int)((N + b - 1) / b);
    int rc = hs_reset_info(h);
    if (rc) return rc;
    const int halfCuRows = hs_half_cu_rows(b);
    const int P = hs_schedule_panels(n);
    // The chain of panels k0 .. k0+P-1 (k0 > 0) runs beside the rest update
    // of step k0 - P, over m = n - k0 - P rows. When that update is half-CU,
    // its POTRFs ask for more LDS than the free half holds (80 KB beside
    // 83.5 KB), so they start only on a CU with no update work-group: beside
    // one, the POTRF's dependent FP64 arithmetic waits behind the update's
    // FP64 MFMAs and takes three to four times as long.
    auto potrf = [&](int k) {
        const int k0 = k / P * P, m = n - k0 - P;
        const size_t minLds = (k0 > 0 && m >= 1 && m <= halfCuRows) ? HALF_CU_POTRF_LDS : 0;
        return hs_factor_potrf<double>(h, A, N, b, hs_block_id(k, k, n), k, minLds);
    };
    auto trsm = [&](int k, int start, int count) {
        return hs_factor_trsm<double>(h, A, b, hs_block_id(k, k, n), k, start, count);
    };
    auto trail = [&](hipStream_t s, int k, int mode, int rowStart, int rowCount, int np) {
        return hs_launch_trailing_on(h, s, A, b, k, rowStart, rowCount, n, mode, np,
                                     mode == 4 && rowCount <= halfCuRows);
    };
    return hs_lookahead_factor(h, n, pt, potrf, trsm, trail);
}

// ---------------------------------------------------------------------------
// Whole factorisation, Cholesky::solve_heterogeneous at share 1.0
// (Cholesky.cpp:446-476): per column k, POTRF of the diagonal block, TRSM of
// the n-k-1 blocks below it, then the SYRK+GEMM trailing update. For b a
// multiple of 64 the look-ahead driver above runs the panel work of step k+1
// underneath the trailing update of step k; other block sizes run the plain
// sequence on one stream. Everything is stream ordered; the only host
// synchronisation is the final read of the pivot status.
// ---------------------------------------------------------------------------
static int cholesky_sequential(hs_handle_t h, double* A, i64 N, int b, PhaseTimer& pt) {
    const int n = (int)((N + b - 1) / b);
    int rc = hs_reset_info(h);
    if (rc) return rc;
    for (int k = 0; k < n; ++k) {
        const i64 id = hs_block_id(k, k, n);
        if ((rc = pt.begin(h->stream, 0, k))) return rc;
        if ((rc = hs_factor_potrf<double>(h, A, N, b, id, k))) return rc;
        if ((rc = pt.end(h->stream, 0))) return rc;
        const int cnt = n - k - 1;
        if (cnt > 0) {
            if ((rc = pt.begin(h->stream, 1, k))) return rc;
            if ((rc = hs_factor_trsm<double>(h, A, b, id, k, k + 1, cnt))) return rc;
            if ((rc = pt.end(h->stream, 1))) return rc;
            if ((rc = pt.begin(h->stream, 2, k))) return rc;
            if ((rc = hs_launch_trailing(h, A, b, k, k + 1, cnt, n, 3))) return rc;
            if ((rc = pt.end(h->stream, 2))) return rc;
        }
    }
    return HS_OK;
}

static int cholesky_run(hs_handle_t h, double* A, i64 N, int b, PhaseTimer& pt) {
    int rc = (b % 64 == 0) ? cholesky_lookahead(h, A, N, b, pt) : cholesky_sequential(h, A, N, b, pt);
    if (rc) {
        hipStreamSynchronize(h->stream2);
        return rc;
    }
    int64_t info = 0;
    if ((rc = hs_get_info(h, &info))) return rc;  // synchronises the handle's stream
    return (int)info;
}

extern "C" int hs_cholesky_f64(hs_handle_t h, double* A, int64_t N, int b) {
    if (!h) return -1;
    if (!A) return -2;
    if (N <= 0) return -3;
    if (b <= 0) return -4;
    PhaseTimer pt;
    return cholesky_run(h, A, N, b, pt);
}

// phase 0 = POTRF launches, 1 = TRSM launches, 2 = trailing-update launches
// on the main stream (the look-ahead column updates run on the panel stream
// and are not counted in phase 2).
extern "C" int hs_cholesky_f64_timed(hs_handle_t h, double* A, int64_t N, int b, double* phase_ms,
                                     int64_t* launches) {
    if (!h) return -1;
    if (!A) return -2;
    if (N <= 0) return -3;
    if (b <= 0) return -4;
    if (!phase_ms) return -5;
    if (!launches) return -6;
    PhaseTimer pt;
    pt.on = true;
    int rc = cholesky_run(h, A, N, b, pt);
    if (rc < 0) return rc;
    int rc2 = pt.collect(phase_ms, launches);
    return rc2 ? rc2 : rc;
}

// Per block column k: the POTRF of diagonal block k, the TRSM of panel k, and
// the trailing-update launches whose first panel is k (host arrays of n
// entries, ms); the reference's per-column record (Cholesky.cpp:261-359).
extern "C" int hs_cholesky_f64_trace(hs_handle_t h, double* A, int64_t N, int b, double* potrf_ms, double* trsm_ms,
                                     double* update_ms) {
    if (!h) return -1;
    if (!A) return -2;
    if (N <= 0) return -3;
    if (b <= 0) return -4;
    if (!potrf_ms) return -5;
    if (!trsm_ms) return -6;
    if (!update_ms) return -7;
    PhaseTimer pt;
    pt.on = true;
    pt.columns = true;
    int rc = cholesky_run(h, A, N, b, pt);
    if (rc < 0) return rc;
    int rc2 = pt.collect_columns((int)((N + b - 1) / b), potrf_ms, trsm_ms, update_ms);
    return rc2 ? rc2 : rc;
}

size_t hs_ws_tri_solve(i64 N, int b) {
    if (!hs_trsv_fast_ok(b)) return 0;
    const int n = (int)((N + b - 1) / b);
    return sizeof(double) * ((size_t)hs_dinv_doubles(n, b) + (size_t)n * b);
}

// the reference's per-block sequence (TriangularSystemSolver.cpp:31-82) for
// block sizes the blocked kernels do not take
static int tri_solve_blockwise(hs_handle_t h, const double* L, double* x, i64 N, int b, int phases) {
    const int n = (int)((N + b - 1) / b);
    int rc;
    if (phases & 1) {
        for (int j = 0; j < n; ++j) {
            const i64 id = hs_block_id(j, j, n);
            if ((rc = hs_launch_trsv_block(h, L, x, N, b, j, id, 0))) return rc;
            if (j < n - 1 && (rc = hs_launch_col_update(h, L, x, b, j + 1, n - (j + 1), j, id, n, 0))) return rc;
        }
    }
    if (phases & 2) {
        for (int j = n - 1; j >= 0; --j) {
            const i64 id = hs_block_id(j, j, n);
            if ((rc = hs_launch_trsv_block(h, L, x, N, b, j, id, 1))) return rc;
            if (j > 0 && (rc = hs_launch_col_update(h, L, x, b, 0, j, j, id, n, 1))) return rc;
        }
    }
    return HS_OK;
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
    if (hs_trsv_fast_ok(b)) {
        // workspace: the inverted diagonal (super-)blocks, then the forward vector
        const size_t nd = (size_t)hs_dinv_doubles(n, b);
        double* Dinv = nullptr;
        if ((rc = hs_work(h, hs_ws_tri_solve(N, b), (void**)&Dinv))) return rc;
        double* z = Dinv + nd;
        if ((rc = hs_diag_inverse<double>(h, L, N, b, Dinv))) return rc;
        return hs_tri_solve_prepared<double>(h, L, Dinv, x, z, N, b);
    }
    return tri_solve_blockwise(h, L, x, N, b, 3);
}

// Prepared factor: the inverted diagonal blocks and 1024-row super-blocks of
// L, built once into a caller buffer, so that repeated solves with the same
// factor skip that step.
extern "C" int hs_tri_prepare_query(int64_t N, int b, size_t* bytes) {
    if (N <= 0) return -1;
    if (b <= 0) return -2;
    if (!bytes) return -3;
    *bytes = hs_trsv_fast_ok(b) ? sizeof(double) * (size_t)hs_dinv_doubles((int)((N + b - 1) / b), b) : 0;
    return HS_OK;
}

extern "C" int hs_tri_prepare_f64(hs_handle_t h, const double* L, int64_t N, int b, double* prep) {
    if (!h) return -1;
    if (!L) return -2;
    if (N <= 0) return -3;
    if (b <= 0) return -4;
    if (!hs_trsv_fast_ok(b)) return HS_OK;  // nothing to prepare: the block-wise solve reads L alone
    if (!prep) return -5;
    return hs_diag_inverse<double>(h, L, N, b, prep);
}

extern "C" int hs_tri_solve_prepared_f64(hs_handle_t h, const double* L, const double* prep, double* x, int64_t N,
                                         int b) {
    if (!h) return -1;
    if (!L) return -2;
    if (N <= 0) return -5;
    if (b <= 0) return -6;
    if (!x) return -4;
    if (!hs_trsv_fast_ok(b)) return tri_solve_blockwise(h, L, x, N, b, 3);
    if (!prep) return -3;
    const int n = (int)((N + b - 1) / b);
    double* z = nullptr;
    int rc = hs_work(h, sizeof(double) * (size_t)n * b, (void**)&z);
    if (rc) return rc;
    return hs_tri_solve_prepared<double>(h, L, prep, x, z, N, b);
}

// One triangular solve in place on x: L x' = x (trans 0) or L^T x' = x
// (trans 1); one half of TriangularSystemSolver::solve
// (TriangularSystemSolver.cpp:31-56 forward, :58-82 backward).
extern "C" int hs_trsv_packed_f64(hs_handle_t h, const double* L, double* x, int64_t N, int b, int trans) {
    if (!h) return -1;
    if (!L) return -2;
    if (!x) return -3;
    if (N <= 0) return -4;
    if (b <= 0) return -5;
    if (trans != 0 && trans != 1) return -6;
    const int n = (int)((N + b - 1) / b);
    int rc;
    if (hs_trsv_fast_ok(b)) {
        const size_t nb = (size_t)n * b;
        const size_t nd = (size_t)hs_dinv_doubles(n, b);
        double* Dinv = nullptr;
        if ((rc = hs_work(h, hs_ws_tri_solve(N, b), (void**)&Dinv))) return rc;
        double* z = Dinv + nd;
        if ((rc = hs_diag_inverse<double>(h, L, N, b, Dinv))) return rc;
        if (trans == 0) {
            if ((rc = hs_tri_solve_prepared<double>(h, L, Dinv, x, z, N, b, 1))) return rc;
            HS_HIP(hipMemcpyAsync(x, z, sizeof(double) * nb, hipMemcpyDeviceToDevice, h->stream));
        } else {
            HS_HIP(hipMemcpyAsync(z, x, sizeof(double) * nb, hipMemcpyDeviceToDevice, h->stream));
            if ((rc = hs_tri_solve_prepared<double>(h, L, Dinv, x, z, N, b, 2))) return rc;
        }
        return HS_OK;
    }
    return tri_solve_blockwise(h, L, x, N, b, trans == 0 ? 1 : 2);
}

// Host-only executor of the look-ahead schedule: each operation becomes 8
// int64 values (kind, stream, a0 .. a5) instead of a launch.
//   0 record(event)        1 wait(event)
//   2 potrf(k)             3 trsm(k, blockStart, blockCount)
//   4 trail(k, mode, rowStart, rowCount, np, reads_planes)
//   5 prep(k0)             (the mixed factor's split of panels k0 .. k0+P-1)
namespace {
struct DryRunSchedule {
    int n, P;
    bool planes;
    std::vector<int64_t> ops;
    void put(int64_t kind, int s, int64_t a0 = 0, int64_t a1 = 0, int64_t a2 = 0, int64_t a3 = 0, int64_t a4 = 0,
             int64_t a5 = 0) {
        for (int64_t v : {kind, (int64_t)s, a0, a1, a2, a3, a4, a5}) ops.push_back(v);
    }
    int record(int ev, int s) { return put(0, s, ev), HS_OK; }
    int wait(int s, int ev) { return put(1, s, ev), HS_OK; }
    int potrf(int k) { return put(2, 2, k), HS_OK; }
    int trsm(int k, int start, int count) { return put(3, 2, k, start, count), HS_OK; }
    int trail(int s, int k, int mode, int rowStart, int rowCount, int np) {
        put(4, s, k, mode, rowStart, rowCount, np, planes && hs_trail_reads_planes(mode, np, P));
        return HS_OK;
    }
    int prep(int k0) {
        if (planes && hs_plane_prep_needed(k0, P, n)) put(5, 2, k0);
        return HS_OK;
    }
    int begin(int, int, int) { return HS_OK; }
    int end(int, int) { return HS_OK; }
};
}  // namespace

// The look-ahead schedule for n block columns and `panels` panels per step
// (0: the FP64 factor's default; clamped to 1 .. n/2 as the factors do), with
// the mixed factor's bf16 plane splits when planes != 0, recorded on the host.
// *count receives the number of operations; at most `capacity` are written.
extern "C" int hs_lookahead_dry_run(int n, int panels, int planes, int64_t* ops, int64_t capacity, int64_t* count) {
    if (n <= 0) return -1;
    if (!ops && capacity > 0) return -4;
    if (!count) return -6;
    DryRunSchedule ex;
    ex.n = n;
    ex.P = panels > 0 ? std::max(1, std::min(panels, n / 2)) : hs_schedule_panels(n);
    ex.planes = planes != 0;
    int rc = hs_lookahead_schedule(ex, n, ex.P);
    if (rc) return rc;
    const int64_t cnt = (int64_t)ex.ops.size() / 8;
    *count = cnt;
    for (int64_t i = 0; i < std::min(cnt, capacity) * 8; ++i) ops[i] = ex.ops[(size_t)i];
    return HS_OK;
}

// The work-group map of one trailing-update launch, replayed on the host with
// the kernels' own decode: for work-group g, out[3g .. 3g+2] = (I, J, tile),
// or (-1, -1, -1) for an idle work-group. band <= 0 and order < 0 take the
// library's defaults for this launch. *count receives the grid size; at most
// `capacity` work-groups are written.
extern "C" int hs_trail_map_dry_run(int n, int k, int rowStart, int rowCount, int mode, int np, int tilesPerBlock,
                                    int band, int order, int32_t* out, int64_t capacity, int64_t* count) {
    if (n <= 0) return -1;
    if (mode < 1 || mode > 6) return -5;
    if (np < 1) return -6;
    if (tilesPerBlock < 1) return -7;
    if (band > 16) return -8;
    if (order > 3) return -9;
    if (!out && capacity > 0) return -10;
    if (!count) return -12;
    int bits, h;
    hs_trail_order(tilesPerBlock, np, &bits, &h);
    if (band > 0) h = band;
    if (order >= 0) bits = order;
    PairMap pm;
    const i64 grid = hs_pair_map_order(k, rowStart, rowCount, n, mode, np, tilesPerBlock, bits, h, &pm);
    *count = grid;
    const i64 w = std::min<i64>(grid, capacity);
    for (i64 g = 0; g < w; ++g) {
        int I = -1, J = -1, tile = -1;
        if (!tile_of_g(pm, tilesPerBlock, g, &I, &J, &tile)) I = J = tile = -1;
        out[3 * g] = I;
        out[3 * g + 1] = J;
        out[3 * g + 2] = tile;
    }
    return HS_OK;
}

// Work of the rest-update launches of the look-ahead schedule the factorisations
// use for N and b (b a multiple of 64): their count and algorithmic flops.
extern "C" int hs_cholesky_schedule_f64(int64_t N, int b, int64_t* rest_launches, double* rest_flops) {
    if (N <= 0) return -1;
    if (b <= 0 || b % 64 != 0) return -2;
    if (!rest_launches) return -3;
    if (!rest_flops) return -4;
    hs_rest_update_work((int)((N + b - 1) / b), b, rest_launches, rest_flops);
    return HS_OK;
}
