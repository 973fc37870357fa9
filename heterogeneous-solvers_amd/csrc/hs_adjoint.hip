// Gaussian-process negative log marginal likelihood and its gradient with
// respect to theta = (log ell, log sf2, log sn2), by a hand-written reverse
// pass through the blocked Cholesky factorisation. There is no counterpart in
// the reference; the forward pass reuses its kernel matrix (MatrixGenerator.cpp:
// 143-153) and the blocked factorisation and solves of this library.
//
//   K = sf2 exp(-|x_i - x_j|^2 / (2 ell^2)) + sn2 I,   K = L L^T
//   z = L^{-1} y,  alpha = L^{-T} z
//   NLL = 1/2 z^T z + sum_i log L_ii + N/2 log(2 pi)
//
// Adjoints are kept in the packed lower-block layout with the lower-storage
// convention: dF = sum_{i >= j} Abar_ij dA_ij, i.e. an off-diagonal entry
// carries the contributions of (i,j) and (j,i) and upper triangles of
// diagonal blocks stay zero. Starting from
//   Lbar = diag(1 / L_ii) - tril(alpha z^T),
// block columns k = n-1 .. 0 are reversed (forward step k: POTRF, TRSM, then
// the trailing update T -= tril(P P^T) with P the panel below L_kk = D):
//   update:  Pbar -= (S + S^T) P,        S = trailing adjoint (lower storage)
//   TRSM:    Abar_panel = Pbar D^{-1},   Dbar -= tril(Abar_panel^T P)
//   POTRF:   M = D^T Dbar, Phi = tril(M) - diag(M)/2,
//            Ssym = D^{-T} Phi D^{-1},   Abar_kk = tril(Ssym + Ssym^T) - diag(Ssym)
// The result is dNLL/dK in lower storage; the gradient contracts it with
// dK/dtheta recomputed from X on the fly:
//   g_ell = sum Abar Kf r^2 / ell^2,  g_sf2 = sum Abar Kf,  g_sn2 = sn2 tr(Abar)
// with Kf = sf2 exp(-r^2 / (2 ell^2)).
//
// The products run on v_mfma_f64_16x16x4f64 in 64x64 output tiles with
// operands staged through LDS; the trailing-adjoint product is split over
// block-column chunks into a workspace and summed in a fixed order, so the
// result does not depend on scheduling.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "hs_internal.h"

typedef double ad4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int AT = 64;       // output tile edge
constexpr int AKT = 16;      // k stage depth
constexpr int ALDP = AT + 16;  // LDS pitch (doubles): lane groups of one read hit disjoint bank halves
constexpr int ATHREADS = 256;

// Stage loaders: each thread moves 4 consecutive doubles of a 64 x 16 (m x k)
// or 16 x 64 (k x n) operand tile into the k-major LDS stage S[k][m].
//   rowk: element (m, k) at p + m * ld + k   (k contiguous)
//   km:   element (m, k) at p + k * ld + m   (m contiguous)
__device__ inline void stage_rowk(double* S, const double* p, int ld, int tid, bool add = false) {
    const int m = tid >> 2, k = (tid & 3) * 4;
    const double2 a = *reinterpret_cast<const double2*>(p + (i64)m * ld + k);
    const double2 c = *reinterpret_cast<const double2*>(p + (i64)m * ld + k + 2);
    const double v[4] = {a.x, a.y, c.x, c.y};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        if (add) S[(k + e) * ALDP + m] += v[e];
        else S[(k + e) * ALDP + m] = v[e];
    }
}
__device__ inline void stage_km(double* S, const double* p, int ld, int tid, bool add = false) {
    const int k = tid >> 4, m = (tid & 15) * 4;
    const double2 a = *reinterpret_cast<const double2*>(p + (i64)k * ld + m);
    const double2 c = *reinterpret_cast<const double2*>(p + (i64)k * ld + m + 2);
    double* s = S + k * ALDP + m;
    if (add) {
        s[0] += a.x;
        s[1] += a.y;
        s[2] += c.x;
        s[3] += c.y;
    } else {
        s[0] = a.x;
        s[1] = a.y;
        s[2] = c.x;
        s[3] = c.y;
    }
}

// One staged k block: 4 waves in 2 x 2, each a 32 x 32 quarter of the tile.
__device__ inline void mma_stage(ad4 (&acc)[2][2], const double* As, const double* Bs, int wave, int lane) {
    const int lr = lane & 15, lq = lane >> 4;
    const int wm0 = (wave >> 1) * 32, wn0 = (wave & 1) * 32;
#pragma unroll
    for (int ks = 0; ks < AKT; ks += 4) {
        double a[2], bv[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = As[(ks + lq) * ALDP + wm0 + i * 16 + lr];
#pragma unroll
        for (int j = 0; j < 2; ++j) bv[j] = Bs[(ks + lq) * ALDP + wn0 + j * 16 + lr];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bv[j], acc[i][j], 0, 0, 0);
    }
}

// Output (row, col) of accumulator register v: row l/16 + 4v, column l%16.
template <class F>
__device__ inline void for_each_out(const ad4 (&acc)[2][2], int wave, int lane, F f) {
    const int lr = lane & 15, lq = lane >> 4;
    const int wm0 = (wave >> 1) * 32, wn0 = (wave & 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) f(wm0 + i * 16 + lq + 4 * v, wn0 + j * 16 + lr, acc[i][j][v]);
}

// Batched C_g = op(A_g) op(B_g) over b x b-sized operands (K a multiple of
// 16, tiles of 64). alay 0: A rowk, 1: A km; B is always km (B[k][n] at
// B + k * ldb + n). mode 0 stores C, mode 1 stores Phi(C) = tril(C) with
// the diagonal halved and zero above it.
struct GemmArgs {
    const double* A;
    i64 strideA;
    int lda, alay;
    const double* B;
    i64 strideB;
    int ldb;
    double* C;
    i64 strideC;
    int ldc;
    int K, tilesN, tilesPer, mode;
};

__global__ __launch_bounds__(ATHREADS) void adj_gemm_kernel(GemmArgs g) {
    __shared__ double As[AKT * ALDP];
    __shared__ double Bs[AKT * ALDP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bi = blockIdx.x / g.tilesPer, tile = blockIdx.x % g.tilesPer;
    const int tm = tile / g.tilesN, tn = tile % g.tilesN;
    const double* A = g.A + bi * g.strideA;
    const double* B = g.B + bi * g.strideB + tn * AT;
    double* C = g.C + bi * g.strideC;
    ad4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) acc[i][0] = acc[i][1] = ad4{0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < g.K; t0 += AKT) {
        if (g.alay == 0) stage_rowk(As, A + (i64)tm * AT * g.lda + t0, g.lda, tid);
        else stage_km(As, A + (i64)t0 * g.lda + tm * AT, g.lda, tid);
        stage_km(Bs, B + (i64)t0 * g.ldb, g.ldb, tid);
        __syncthreads();
        mma_stage(acc, As, Bs, wave, lane);
        __syncthreads();
    }
    const int r0 = tm * AT, c0 = tn * AT;
    const int mode = g.mode;
    for_each_out(acc, wave, lane, [&](int r, int c, double v) {
        const int gr = r0 + r, gc = c0 + c;
        if (mode == 1) v = (gr > gc) ? v : (gr == gc ? 0.5 * v : 0.0);
        C[(i64)gr * g.ldc + gc] = v;
    });
}

// adj_gemm_kernel (row-k A, mode 0) with every column tile of a 64-row strip
// in one work-group: each tile accumulates over k in the same order, and the
// whole A strip is read before any C element is written, so C may be A (the
// panel adjoint times D^{-1} in place, without a copy back).
__global__ __launch_bounds__(ATHREADS) void adj_gemm_strip_kernel(GemmArgs g) {
    constexpr int MAXN = 4;  // b <= 256
    __shared__ double As[AKT * ALDP];
    __shared__ double Bs[AKT * ALDP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tilesM = g.tilesPer / g.tilesN;
    const int bi = blockIdx.x / tilesM, tm = blockIdx.x % tilesM;
    const double* A = g.A + bi * g.strideA + (i64)tm * AT * g.lda;
    const double* B = g.B + bi * g.strideB;
    double* C = g.C + bi * g.strideC;
    ad4 acc[MAXN][2][2];
#pragma unroll
    for (int t = 0; t < MAXN; ++t)
#pragma unroll
        for (int i = 0; i < 2; ++i) acc[t][i][0] = acc[t][i][1] = ad4{0.0, 0.0, 0.0, 0.0};
    for (int t0 = 0; t0 < g.K; t0 += AKT) {
        stage_rowk(As, A + t0, g.lda, tid);
#pragma unroll
        for (int tn = 0; tn < MAXN; ++tn) {
            if (tn < g.tilesN) {
                stage_km(Bs, B + (i64)t0 * g.ldb + tn * AT, g.ldb, tid);
                __syncthreads();
                mma_stage(acc[tn], As, Bs, wave, lane);
                __syncthreads();
            }
        }
    }
    const int r0 = tm * AT;
#pragma unroll
    for (int tn = 0; tn < MAXN; ++tn) {
        if (tn < g.tilesN)
            for_each_out(acc[tn], wave, lane,
                         [&](int r, int c, double v) { C[(i64)(r0 + r) * g.ldc + tn * AT + c] = v; });
    }
}

// Trailing-adjoint product of reverse step k, one output tile of one panel
// block I and one chunk of block columns J in [Ja, Jb):
//   W[chunk][I - k - 1] = sum_J Sfull_IJ P_J,  Sfull_IJ = S_IJ (J < I),
//   S_JI^T (J > I), S_II + S_II^T (J == I); S = Abar, P_J = L block (J, k).
__global__ __launch_bounds__(ATHREADS) void adj_symm_kernel(const double* __restrict__ Ab, const double* __restrict__ L,
                                                            double* __restrict__ W, int b, int n, int k, int m,
                                                            int chunk, int tilesN) {
    __shared__ double As[AKT * ALDP];
    __shared__ double Bs[AKT * ALDP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tilesPer = tilesN * tilesN;
    const int tile = blockIdx.x % tilesPer;
    const int rest = blockIdx.x / tilesPer;
    const int ip = rest % m;      // I - k - 1
    const int c = rest / m;       // chunk index
    const int I = k + 1 + ip;
    const int Ja = k + 1 + c * chunk, Jb = std::min(n, Ja + chunk);
    const int tm = tile / tilesN, tn = tile % tilesN;
    const i64 bb = (i64)b * b;
    ad4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) acc[i][0] = acc[i][1] = ad4{0.0, 0.0, 0.0, 0.0};
    for (int J = Ja; J < Jb; ++J) {
        const double* P = L + hs_block_id(J, k, n) * bb + tn * AT;
        const double* Sij = (J <= I) ? Ab + hs_block_id(I, J, n) * bb : nullptr;
        const double* Sji = (J >= I) ? Ab + hs_block_id(J, I, n) * bb : nullptr;
        for (int t0 = 0; t0 < b; t0 += AKT) {
            if (J < I) {
                stage_rowk(As, Sij + (i64)tm * AT * b + t0, b, tid);
            } else if (J > I) {
                stage_km(As, Sji + (i64)t0 * b + tm * AT, b, tid);
            } else {
                stage_rowk(As, Sij + (i64)tm * AT * b + t0, b, tid);
                __syncthreads();
                stage_km(As, Sji + (i64)t0 * b + tm * AT, b, tid, true);
            }
            stage_km(Bs, P + (i64)t0 * b, b, tid);
            __syncthreads();
            mma_stage(acc, As, Bs, wave, lane);
            __syncthreads();
        }
    }
    double* Wt = W + ((i64)c * m + ip) * bb;
    const int r0 = tm * AT, c0 = tn * AT;
    for_each_out(acc, wave, lane, [&](int r, int cc, double v) { Wt[(i64)(r0 + r) * b + c0 + cc] = v; });
}

// The same product in 128 x 128 tiles for b % 128 == 0: the trailing-update
// kernel's structure (2 x 2 waves of 64 x 64 f64 MFMA accumulators, 16-deep
// k stages double-buffered in LDS with the next stage's global loads in
// flight), with the three A layouts of Sfull (S_IJ row-major, S_JI^T read
// column-wise, S_II + S_II^T) and the panel block P_J read row-wise as B.
// The partial tiles are written once and read once by the summing pass:
// non-temporal stores keep them from displacing operands in the cache.
constexpr int ST = 128;               // tile edge
constexpr int SKT = 16;               // k stage depth
constexpr int SLDP = ST + 4;          // LDS pitch (doubles)
constexpr int SSTAGE = SKT * SLDP;

// One launch covers the panel blocks I = I0 .. I0 + nrows - 1 and the block
// columns J0 .. J1 - 1 in chunks; chunk c of row I goes to partial slot
// slot0 + c: W[(slot0 + c) m + I - k - 1]. A launch may carry a second such
// range (nrows2 > 0) after the first one's nrows x nch work-groups.
struct SymmPart {
    int I0, nrows, J0, J1, chunk, slot0;
    int nch = 0, I02 = 0, nrows2 = 0, J02 = 0, J12 = 0, chunk2 = 1, slot02 = 0;
};

__global__ __launch_bounds__(256) void adj_symm128_kernel(const double* __restrict__ Ab, const double* __restrict__ L,
                                                          double* __restrict__ W, int b, int n, int k, int m,
                                                          SymmPart sp, int tilesN) {
    __shared__ double As[2 * SSTAGE];
    __shared__ double Bs[2 * SSTAGE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tilesPer = tilesN * tilesN;
    const int tile = blockIdx.x % tilesPer;
    int rest = blockIdx.x / tilesPer;
    int I, c, Ja, Jb, slot;
    if (sp.nrows2 > 0 && rest >= sp.nrows * sp.nch) {
        rest -= sp.nrows * sp.nch;
        I = sp.I02 + rest % sp.nrows2;
        c = rest / sp.nrows2;
        Ja = sp.J02 + c * sp.chunk2;
        Jb = std::min(sp.J12, Ja + sp.chunk2);
        slot = sp.slot02 + c;
    } else {
        I = sp.I0 + rest % sp.nrows;
        c = rest / sp.nrows;
        Ja = sp.J0 + c * sp.chunk;
        Jb = std::min(sp.J1, Ja + sp.chunk);
        slot = sp.slot0 + c;
    }
    const int ip = I - k - 1;
    const int tm = tile / tilesN, tn = tile % tilesN;
    const i64 bb = (i64)b * b;
    const int nkb = b / SKT;                  // stages per J block
    const int nst = (Jb - Ja) * nkb;          // stages of this work-group
    // thread -> staging positions: row-major A (8 consecutive k of one row)
    // and column-wise A / B (8 consecutive m of one k row)
    const int lrow = tid >> 1, lk = (tid & 1) * 8;
    const int kr = tid >> 4, m0 = (tid & 15) * 8;
    double2 ra[4], ra2[4], rb[4];
    int sJ = 0;  // J of the stage held in registers
    auto gload = [&](int st) {
        const int J = Ja + st / nkb, t0 = (st % nkb) * SKT;
        sJ = J;
        if (J <= I) {
            const double* p = Ab + hs_block_id(I, J, n) * bb + (i64)(tm * ST + lrow) * b + t0 + lk;
#pragma unroll
            for (int v = 0; v < 4; ++v) ra[v] = *reinterpret_cast<const double2*>(p + 2 * v);
        }
        if (J >= I) {
            const double* p = Ab + hs_block_id(J, I, n) * bb + (i64)(t0 + kr) * b + tm * ST + m0;
#pragma unroll
            for (int v = 0; v < 4; ++v) ra2[v] = *reinterpret_cast<const double2*>(p + 2 * v);
        }
        const double* pb = L + hs_block_id(J, k, n) * bb + (i64)(t0 + kr) * b + tn * ST + m0;
#pragma unroll
        for (int v = 0; v < 4; ++v) rb[v] = *reinterpret_cast<const double2*>(pb + 2 * v);
    };
    auto sstore = [&](int buf) {
        double* as = As + buf * SSTAGE;
        double* bs = Bs + buf * SSTAGE;
        if (sJ < I) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                as[(lk + 2 * v) * SLDP + lrow] = ra[v].x;
                as[(lk + 2 * v + 1) * SLDP + lrow] = ra[v].y;
            }
        } else if (sJ > I) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                as[kr * SLDP + m0 + 2 * v] = ra2[v].x;
                as[kr * SLDP + m0 + 2 * v + 1] = ra2[v].y;
            }
        } else {
            // S_II + S_II^T: the row-major part first, then the transposed
            // part added after a barrier (uniform: sJ is the same in the group)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                as[(lk + 2 * v) * SLDP + lrow] = ra[v].x;
                as[(lk + 2 * v + 1) * SLDP + lrow] = ra[v].y;
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                as[kr * SLDP + m0 + 2 * v] += ra2[v].x;
                as[kr * SLDP + m0 + 2 * v + 1] += ra2[v].y;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            bs[kr * SLDP + m0 + 2 * v] = rb[v].x;
            bs[kr * SLDP + m0 + 2 * v + 1] = rb[v].y;
        }
    };
    ad4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = ad4{0.0, 0.0, 0.0, 0.0};
    const int lr = lane & 15, lq = lane >> 4;
    const int wm0 = (wave >> 1) * 64, wn0 = (wave & 1) * 64;
    gload(0);
    sstore(0);
    if (nst > 1) gload(1);
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const double* as = As + (st & 1) * SSTAGE;
        const double* bs = Bs + (st & 1) * SSTAGE;
#pragma unroll
        for (int ks = 0; ks < SKT; ks += 4) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = as[(ks + lq) * SLDP + wm0 + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = bs[(ks + lq) * SLDP + wn0 + j * 16 + lr];
            // the other buffer was last read in stage st-1, which every wave
            // left at the barrier: fill it between this stage's MFMAs
            if (ks == 0 && st + 1 < nst) {
                sstore((st + 1) & 1);
                if (st + 2 < nst) gload(st + 2);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // output: register v of lane l is row l/16 + 4v, column l%16 of its tile
    double* Wt = W + ((i64)slot * m + ip) * bb + (i64)(tm * ST) * b + tn * ST;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                __builtin_nontemporal_store(acc[i][j][v], Wt + (i64)(wm0 + i * 16 + lq + 4 * v) * b + wn0 + j * 16 + lr);
}

// s + sum_{c < nch} W[c * stride] in order c = 0, 1, ..., with the loads of
// 8 consecutive c issued together (one load at a time left the sums of the
// reverse sweep latency-bound at 1-2 TB/s). The partials are read once.
__device__ inline double sum_slots(const double* __restrict__ W, i64 stride, int nch) {
    double s = 0.0;
    int c = 0;
    for (; c + 8 <= nch; c += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = __builtin_nontemporal_load(W + (i64)(c + u) * stride);
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c < nch; ++c) s += __builtin_nontemporal_load(W + (i64)c * stride);
    return s;
}

// dst[e] -= sum_{c < nch} W[c * stride + e], e < count (fixed order)
__global__ void adj_sub_partials_kernel(double* __restrict__ dst, const double* __restrict__ W, i64 count, i64 stride,
                                        int nch) {
    for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (i64)gridDim.x * blockDim.x)
        dst[e] -= sum_slots(W + e, stride, nch);
}

__global__ void adj_copy_kernel(double* __restrict__ dst, const double* __restrict__ src, i64 count) {
    for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (i64)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

// Dbar (b x b, lower) -= tril(sum_c Wg[c]) with Wg[c] = (Abar_panel^T P)
// partials, summed in order c = 0 .. nch-1; thread t takes the t-th element
// of the lower triangle in row order (row r starts at r (r + 1) / 2)
__global__ void adj_gram_sub_kernel(double* __restrict__ Dbar, const double* __restrict__ Wg, int b, int nch) {
    const i64 bb = (i64)b * b, tri = (i64)b * (b + 1) / 2;
    for (i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x; t < tri; t += (i64)gridDim.x * blockDim.x) {
        int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((i64)r * (r + 1) / 2 > t) --r;
        while ((i64)(r + 1) * (r + 2) / 2 <= t) ++r;
        const int c = (int)(t - (i64)r * (r + 1) / 2);
        const i64 e = (i64)r * b + c;
        Dbar[e] -= sum_slots(Wg + e, bb, nch);
    }
}

// Abar_kk = tril(Ssym + Ssym^T) - diag(Ssym), zero above the diagonal
__global__ void adj_symmetrize_kernel(double* __restrict__ Akk, const double* __restrict__ Ssym, int b) {
    const i64 bb = (i64)b * b;
    for (i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x; e < bb; e += (i64)gridDim.x * blockDim.x) {
        const int r = (int)(e / b), c = (int)(e % b);
        double v = 0.0;
        if (r > c) v = Ssym[e] + Ssym[(i64)c * b + r];
        else if (r == c) v = Ssym[e];
        Akk[e] = v;
    }
}

// Lbar = diag(1 / L_ii) - tril(alpha z^T) over the valid entries, zero
// elsewhere; one work-group per stored block.
__global__ void adj_init_kernel(double* __restrict__ Ab, const double* __restrict__ L, const double* __restrict__ alpha,
                                const double* __restrict__ z, i64 N, int b, int n) {
    const int I = blockIdx.x / n, J = blockIdx.x % n;
    if (I < J) return;
    const i64 bb = (i64)b * b;
    double* blk = Ab + hs_block_id(I, J, n) * bb;
    const double* Ld = L + hs_block_id(I, I, n) * bb;
    for (int e = threadIdx.x; e < b * b; e += blockDim.x) {
        const int r = e / b, c = e % b;
        const i64 gi = (i64)I * b + r, gj = (i64)J * b + c;
        double v = 0.0;
        if (gi < N && gj < N && gi >= gj) {
            v = -alpha[gi] * z[gj];
            if (gi == gj) v += 1.0 / Ld[(i64)r * b + r];
        }
        blk[e] = v;
    }
}

// Per stored block: sums of Abar Kf r^2 / ell^2, Abar Kf and the diagonal of
// Abar over the valid lower entries, with Kf recomputed from X (staged in
// LDS; b * d <= 4096). out[block][3], reduced in a fixed order afterwards.
__global__ __launch_bounds__(256) void adj_contract_kernel(const double* __restrict__ Ab,
                                                           const double* __restrict__ X, i64 N, int d, int b, int n,
                                                           double inv_ell2, double sf2, double* __restrict__ out) {
    extern __shared__ double xs[];  // [2][b][d]
    __shared__ double red[3][4];
    const int I = blockIdx.x / n, J = blockIdx.x % n;
    const int tid = threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (I >= J) {
        double* xi = xs;
        double* xj = xs + (size_t)b * d;
        for (int e = tid; e < b * d; e += 256) {
            const int r = e / d, f = e % d;
            const i64 gi = (i64)I * b + r, gj = (i64)J * b + r;
            xi[e] = gi < N ? X[gi * d + f] : 0.0;
            xj[e] = gj < N ? X[gj * d + f] : 0.0;
        }
        __syncthreads();
        const double* blk = Ab + hs_block_id(I, J, n) * (i64)b * b;
        for (int e = tid; e < b * b; e += 256) {
            const int r = e / b, c = e % b;
            const i64 gi = (i64)I * b + r, gj = (i64)J * b + c;
            if (gi >= N || gj >= N || gi < gj) continue;
            const double a = blk[e];
            double dist = 0.0;
            for (int f = 0; f < d; ++f) {
                const double t = xi[r * d + f] - xj[c * d + f];
                dist += t * t;
            }
            const double kf = sf2 * exp(-0.5 * inv_ell2 * dist);
            s0 += a * kf * dist * inv_ell2;
            s1 += a * kf;
            if (gi == gj) s2 += a;
        }
    }
    // wave sums, then the 4 waves in order
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_down(s0, off);
        s1 += __shfl_down(s1, off);
        s2 += __shfl_down(s2, off);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = s0;
        red[1][tid >> 6] = s1;
        red[2][tid >> 6] = s2;
    }
    __syncthreads();
    if (tid < 3) out[(i64)blockIdx.x * 3 + tid] = red[tid][0] + red[tid][1] + red[tid][2] + red[tid][3];
}

// Fixed-order sum of count triples; out[3].
__global__ void adj_sum3_kernel(const double* __restrict__ parts, i64 count, double* __restrict__ out) {
    __shared__ double red[3][256];
    const int tid = threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    for (i64 q = tid; q < count; q += 256)
        for (int j = 0; j < 3; ++j) s[j] += parts[q * 3 + j];
    for (int j = 0; j < 3; ++j) red[j][tid] = s[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int j = 0; j < 3; ++j) red[j][tid] += red[j][tid + w];
        __syncthreads();
    }
    if (tid < 3) out[tid] = red[tid][0];
}

// out[0] = 1/2 z^T z + sum_{i < N} log L_ii (one work-group)
__global__ void adj_nll_kernel(const double* __restrict__ z, const double* __restrict__ L, i64 N, int b, int n,
                               double* __restrict__ out) {
    __shared__ double red[256];
    const int tid = threadIdx.x;
    double s = 0.0;
    for (i64 i = tid; i < N; i += 256) {
        const int I = (int)(i / b), r = (int)(i % b);
        const double l = L[hs_block_id(I, I, n) * (i64)b * b + (i64)r * b + r];
        s += 0.5 * z[i] * z[i] + log(l);
    }
    red[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) out[0] = red[0];
}

inline unsigned grid_for(i64 count) { return (unsigned)std::max<i64>(1, std::min<i64>((count + 255) / 256, 8192)); }

int launch_gemm(hs_handle_t h, const GemmArgs& g, int batch, int tilesM) {
    const i64 grid = (i64)batch * tilesM * g.tilesN;
    adj_gemm_kernel<<<dim3((unsigned)grid), dim3(ATHREADS), 0, h->stream>>>(g);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// chunk of block columns per work-group of the trailing-adjoint product, so
// that one launch has about 2048 work-groups
inline int symm_chunk(int m, int tiles) {
    const int nch = std::max(1, std::min(m, (2048 + m * tiles - 1) / (m * tiles)));
    return (m + nch - 1) / nch;
}

// Partial slots of reverse step k on the 128-tile path (m = n - k - 1 panel
// blocks): slot 0 holds the block column J = k + 1 for every row, slots
// 1 .. bulk_chunks(m) the chunks of J = k + 2 .. n - 1.
inline int bulk_chunk(int m, int stiles) { return symm_chunk(std::max(1, m - 1), stiles); }
inline int bulk_chunks(int m, int stiles) {
    if (m < 2) return 0;
    const int ch = bulk_chunk(m, stiles);
    return (m - 1 + ch - 1) / ch;
}

// Workspace layout (doubles) of one gradient evaluation.
struct AdjWork {
    i64 dinv, v, z, zs, nll, parts, grad, wp, wp_half, wg, t1, t2, t3, total;
};

AdjWork adj_layout(int n, int b) {
    const i64 bb = (i64)b * b, nb = (i64)n * b;
    const int tiles = (b % ST == 0) ? (b / ST) * (b / ST) : (b / AT) * (b / AT);
    // partials of the trailing-adjoint product: one buffer of nch x m blocks
    // (64-tile path), or two alternating buffers of (1 + bulk chunks) x m
    // blocks (128-tile path)
    i64 wp = 0;
    for (int k = 0; k + 1 < n; ++k) {
        const int m = n - k - 1;
        if (b % ST == 0) {
            wp = std::max(wp, (1 + (i64)bulk_chunks(m, tiles)) * m * bb);
        } else {
            const int ch = symm_chunk(m, tiles);
            const i64 nch = (m + ch - 1) / ch;
            wp = std::max(wp, nch * m * bb);
        }
    }
    const i64 wpHalf = (wp + 31) / 32 * 32;
    if (b % ST == 0) wp = 2 * wpHalf;
    AdjWork w;
    i64 o = 0;
    auto take = [&](i64 cnt) {
        const i64 at = o;
        o += (cnt + 31) / 32 * 32;
        return at;
    };
    w.dinv = take(hs_dinv_doubles(n, b));
    w.v = take(nb);
    w.z = take(nb);
    w.zs = take(nb);
    w.nll = take(4);
    w.parts = take(3 * (i64)n * n);
    w.grad = take(4);
    w.wp = take(std::max<i64>(wp, 1));
    w.wp_half = wpHalf;
    w.wg = take((i64)std::max(1, n - 1) * bb);
    w.t1 = take(bb);
    w.t2 = take(bb);
    w.t3 = take(bb);
    w.total = o;
    return w;
}

int check_args(hs_handle_t h, const double* X, int64_t N, int d, int b, double ell, double sf2, double sn2,
               const double* y, double* K) {
    if (!h) return -1;
    if (!X) return -2;
    if (N <= 0) return -3;
    if (d <= 0) return -4;
    if (b % 64 != 0 || b > 256 || b * d > 4096) {
        hs_set_error("GP NLL: matrix_bsz must be 64, 128 or 256 and matrix_bsz * d <= 4096");
        return -5;
    }
    if (!(ell > 0.0)) return -6;
    if (!(sf2 > 0.0)) return -7;
    if (!(sn2 >= 0.0)) return -8;
    if (!y) return -9;
    if (!K) return -10;
    return HS_OK;
}

// Assembly, factorisation and both solves; leaves L in K, z in w.zs, alpha in
// w.v, D^{-T} / D^{-1} in w.dinv and the NLL (without N/2 log 2 pi) in w.nll.
int forward(hs_handle_t h, const double* X, i64 N, int d, int b, double ell, double sf2, double sn2, const double* y,
            double* K, double* ws, const AdjWork& w) {
    const int n = (int)((N + b - 1) / b);
    const i64 nb = (i64)n * b;
    int rc = hs_rbf_assemble_f64(h, X, N, d, b, ell, sf2, sn2, K);
    if (rc) return rc;
    rc = hs_cholesky_f64(h, K, N, b);
    if (rc) return rc;  // > 0: K is not positive definite at that row
    if ((rc = hs_diag_inverse<double>(h, K, N, b, ws + w.dinv))) return rc;
    if ((rc = hs_memset_zero(h, ws + w.v, sizeof(double) * nb))) return rc;
    HS_HIP(hipMemcpyAsync(ws + w.v, y, sizeof(double) * N, hipMemcpyDeviceToDevice, h->stream));
    if ((rc = hs_tri_solve_prepared<double>(h, K, ws + w.dinv, ws + w.v, ws + w.z, N, b, 1))) return rc;
    adj_copy_kernel<<<dim3(grid_for(nb)), dim3(256), 0, h->stream>>>(ws + w.zs, ws + w.z, nb);
    HS_LAUNCH_CHECK();
    if ((rc = hs_tri_solve_prepared<double>(h, K, ws + w.dinv, ws + w.v, ws + w.z, N, b, 2))) return rc;
    adj_nll_kernel<<<1, 256, 0, h->stream>>>(ws + w.zs, K, N, b, n, ws + w.nll);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// The per-step tail of reverse step k on the handle's stream, after the
// trailing-adjoint partials of every panel block are in W (nch of them):
// Pbar -= sum W, the TRSM and POTRF reversals.
int reverse_tail(hs_handle_t h, const double* L, double* Ab, int b, int n, int k, int nch, const double* W,
                 double* ws, const AdjWork& w) {
    const i64 bb = (i64)b * b;
    const int tilesN = b / AT, tiles = tilesN * tilesN;
    const int m = n - k - 1;
    const double* Dup = ws + w.dinv;                // D^{-T} per diagonal block
    const double* Dlo = ws + w.dinv + (i64)n * bb;  // D^{-1}
    double* Akk = Ab + hs_block_id(k, k, n) * bb;
    const double* D = L + hs_block_id(k, k, n) * bb;
    int rc;
    if (m > 0) {
        double* Pbar = Ab + hs_block_id(k + 1, k, n) * bb;     // panel adjoint, m blocks
        const double* P = L + hs_block_id(k + 1, k, n) * bb;   // panel, m blocks
        adj_sub_partials_kernel<<<dim3(grid_for(m * bb)), dim3(256), 0, h->stream>>>(Pbar, W, m * bb, m * bb, nch);
        HS_LAUNCH_CHECK();
        // Abar_panel = Pbar D^{-1}
        // (in place: one work-group per 64-row strip)
        GemmArgs g1{Pbar, bb, b, 0, Dlo + (i64)k * bb, 0, b, Pbar, bb, b, b, tilesN, tiles, 0};
        adj_gemm_strip_kernel<<<dim3((unsigned)((i64)m * tilesN)), dim3(ATHREADS), 0, h->stream>>>(g1);
        HS_LAUNCH_CHECK();
        // Dbar -= tril(Abar_panel^T P): one partial per panel block in a
        // single batched launch, summed in a fixed order
        GemmArgs g2{Pbar, bb, b, 1, P, bb, b, ws + w.wg, bb, b, b, tilesN, tiles, 0};
        if ((rc = launch_gemm(h, g2, m, tilesN))) return rc;
        adj_gram_sub_kernel<<<dim3(grid_for((i64)b * (b + 1) / 2)), dim3(256), 0, h->stream>>>(Akk, ws + w.wg, b, m);
        HS_LAUNCH_CHECK();
    }
    // POTRF reverse on block k
    GemmArgs m1{D, 0, b, 1, Akk, 0, b, ws + w.t1, 0, b, b, tilesN, tiles, 1};   // Phi(D^T Dbar)
    if ((rc = launch_gemm(h, m1, 1, tilesN))) return rc;
    GemmArgs m2{Dup + (i64)k * bb, 0, b, 0, ws + w.t1, 0, b, ws + w.t2, 0, b, b, tilesN, tiles, 0};  // D^{-T} Phi
    if ((rc = launch_gemm(h, m2, 1, tilesN))) return rc;
    GemmArgs m3{ws + w.t2, 0, b, 0, Dlo + (i64)k * bb, 0, b, ws + w.t3, 0, b, b, tilesN, tiles, 0};  // .. D^{-1}
    if ((rc = launch_gemm(h, m3, 1, tilesN))) return rc;
    adj_symmetrize_kernel<<<dim3(grid_for(bb)), dim3(256), 0, h->stream>>>(Akk, ws + w.t3, b);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

int reverse(hs_handle_t h, const double* L, double* Ab, i64 N, int b, double* ws, const AdjWork& w) {
    const int n = (int)((N + b - 1) / b);
    adj_init_kernel<<<dim3((unsigned)((i64)n * n)), dim3(256), 0, h->stream>>>(Ab, L, ws + w.v, ws + w.zs, N, b, n);
    HS_LAUNCH_CHECK();
    int rc;
    if (b % ST != 0) {
        // 64-tile product, one launch per step over every J, then the tail
        const int tilesN = b / AT, tiles = tilesN * tilesN;
        for (int k = n - 1; k >= 0; --k) {
            const int m = n - k - 1;
            int nch = 0;
            if (m > 0) {
                const int ch = symm_chunk(m, tiles);
                nch = (m + ch - 1) / ch;
                adj_symm_kernel<<<dim3((unsigned)((i64)nch * m * tiles)), dim3(ATHREADS), 0, h->stream>>>(
                    Ab, L, ws + w.wp, b, n, k, m, ch, tilesN);
                HS_LAUNCH_CHECK();
            }
            if ((rc = reverse_tail(h, L, Ab, b, n, k, nch, ws + w.wp, ws, w))) return rc;
        }
        return HS_OK;
    }
    // 128-tile path with a look-ahead of one step on the panel stream S2.
    // Reverse step k needs the trailing adjoint S over the blocks (I, J),
    // I, J > k; column k + 1 of it is final only when step k + 1 ends, the
    // rest (I, J >= k + 2) already when step k + 2 ends. So the bulk of step
    // k's product (rows and columns >= k + 2) runs on S2 beside step k + 1,
    // and step k itself adds only block column k + 1 (all rows) and block
    // row k + 1 (columns >= k + 2) before its tail. Events: C[j] = step j
    // done (on S1), B[k] = bulk of step k done (on S2). The partial buffers
    // alternate by step: bulk k - 2 starts after step k's tail has summed
    // them. Step k's bulk reads only columns >= k + 2 while step k + 1
    // writes column k + 1 and block (k + 1, k + 1).
    const int stn = b / ST, stiles = stn * stn;
    if ((rc = hs_ensure_events(h, 2 * (size_t)n + 1))) return rc;
    hipEvent_t* C = h->events.data();
    hipEvent_t* B = h->events.data() + n;
    hipEvent_t start = h->events[2 * (size_t)n];
    double* Wb[2] = {ws + w.wp, ws + w.wp + w.wp_half};
    hipStream_t s1 = h->stream, s2 = h->stream2;
    HS_HIP(hipEventRecord(start, s1));
    HS_HIP(hipStreamWaitEvent(s2, start, 0));
    // sp.nch work-groups per row of the first range, nch2 per row of the
    // second (if any)
    auto launch = [&](hipStream_t st, double* W, int k, const SymmPart& sp, int nch2) -> int {
        const i64 units = (i64)sp.nrows * sp.nch + (sp.nrows2 > 0 ? (i64)sp.nrows2 * nch2 : 0);
        if (units <= 0) return HS_OK;
        const int m = n - k - 1;
        adj_symm128_kernel<<<dim3((unsigned)(units * stiles)), dim3(256), 0, st>>>(Ab, L, W, b, n, k, m, sp, stn);
        HS_LAUNCH_CHECK();
        return HS_OK;
    };
    for (int j = n - 1; j >= 0; --j) {
        // bulk of step j - 1 on S2, beside step j
        const int kb = j - 1, mb = n - kb - 1;
        if (kb >= 0 && mb >= 2) {
            if (j + 1 <= n - 1) HS_HIP(hipStreamWaitEvent(s2, C[j + 1], 0));
            SymmPart sp{kb + 2, mb - 1, kb + 2, n, bulk_chunk(mb, stiles), 1};
            sp.nch = bulk_chunks(mb, stiles);
            if ((rc = launch(s2, Wb[kb & 1], kb, sp, 0))) return rc;
            HS_HIP(hipEventRecord(B[kb], s2));
        }
        // step j on S1
        const int m = n - j - 1;
        const int nchB = bulk_chunks(m, stiles);
        if (m >= 1) {
            // block column j + 1 (every row, slot 0) and block row j + 1
            // (columns j + 2 .., slots 1 .. nchB) in one launch
            SymmPart thin{j + 1, m, j + 1, j + 2, 1, 0};
            thin.nch = 1;
            if (m >= 2) {
                thin.I02 = j + 1;
                thin.nrows2 = 1;
                thin.J02 = j + 2;
                thin.J12 = n;
                thin.chunk2 = bulk_chunk(m, stiles);
                thin.slot02 = 1;
            }
            if ((rc = launch(s1, Wb[j & 1], j, thin, nchB))) return rc;
            if (m >= 2) HS_HIP(hipStreamWaitEvent(s1, B[j], 0));
        }
        if ((rc = reverse_tail(h, L, Ab, b, n, j, 1 + nchB, Wb[j & 1], ws, w))) return rc;
        HS_HIP(hipEventRecord(C[j], s1));
    }
    return HS_OK;
}

}  // namespace

size_t hs_ws_adjoint(i64 N, int b) {
    if (b % 64 != 0 || b > 256) return 0;  // block sizes the adjoint takes
    return sizeof(double) * (size_t)adj_layout((int)((N + b - 1) / b), b).total;
}

extern "C" int hs_gp_nll_f64(hs_handle_t h, const double* X, int64_t N, int d, int b, double ell, double sf2,
                             double sn2, const double* y, double* K, double* nll) {
    int rc = check_args(h, X, N, d, b, ell, sf2, sn2, y, K);
    if (rc) return rc;
    if (!nll) return -11;
    const int n = (int)((N + b - 1) / b);
    const AdjWork w = adj_layout(n, b);
    double* ws = nullptr;
    if ((rc = hs_work(h, sizeof(double) * (size_t)w.total, (void**)&ws))) return rc;
    if ((rc = forward(h, X, N, d, b, ell, sf2, sn2, y, K, ws, w))) return rc;
    double v = 0.0;
    HS_HIP(hipMemcpyAsync(&v, ws + w.nll, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HS_HIP(hipStreamSynchronize(h->stream));
    *nll = v + 0.5 * (double)N * std::log(2.0 * M_PI);
    return HS_OK;
}

extern "C" int hs_gp_nll_grad_f64(hs_handle_t h, const double* X, int64_t N, int d, int b, double ell, double sf2,
                                  double sn2, const double* y, double* K, double* Kbar, double* nll, double* grad) {
    int rc = check_args(h, X, N, d, b, ell, sf2, sn2, y, K);
    if (rc) return rc;
    if (!Kbar) return -11;
    if (!nll) return -12;
    if (!grad) return -13;
    const int n = (int)((N + b - 1) / b);
    const AdjWork w = adj_layout(n, b);
    double* ws = nullptr;
    if ((rc = hs_work(h, sizeof(double) * (size_t)w.total, (void**)&ws))) return rc;
    if ((rc = forward(h, X, N, d, b, ell, sf2, sn2, y, K, ws, w))) return rc;
    if ((rc = reverse(h, K, Kbar, N, b, ws, w))) {
        hipStreamSynchronize(h->stream2);  // no panel-stream work outlives a failed call
        return rc;
    }
    const size_t lds = sizeof(double) * 2 * (size_t)b * d;
    if (lds > 48 * 1024)
        HS_HIP(hipFuncSetAttribute((const void*)adj_contract_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    adj_contract_kernel<<<dim3((unsigned)((i64)n * n)), dim3(256), lds, h->stream>>>(
        Kbar, X, N, d, b, n, 1.0 / (ell * ell), sf2, ws + w.parts);
    HS_LAUNCH_CHECK();
    adj_sum3_kernel<<<1, 256, 0, h->stream>>>(ws + w.parts, (i64)n * n, ws + w.grad);
    HS_LAUNCH_CHECK();
    double hv[4];
    HS_HIP(hipMemcpyAsync(hv, ws + w.nll, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HS_HIP(hipMemcpyAsync(hv + 1, ws + w.grad, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HS_HIP(hipStreamSynchronize(h->stream));
    *nll = hv[0] + 0.5 * (double)N * std::log(2.0 * M_PI);
    grad[0] = hv[1];
    grad[1] = hv[2];
    grad[2] = hv[3] * sn2;
    return HS_OK;
}
