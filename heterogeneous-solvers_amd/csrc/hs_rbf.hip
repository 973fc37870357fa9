// On-device RBF kernel-matrix assembly into the packed block layout.
// Replaces MatrixGenerator::generateSPDMatrix (MatrixGenerator.cpp:73-162),
// which submits one b x b CPU-queue kernel per stored block and waits per
// block row, with one launch of one work-group per stored block.
//
// Per entry, the reference's order: dist = sum_k (x_ik - x_jk)^2 in feature
// order, then sf2 * exp(-0.5/ell^2 * dist), plus sn2 on the global diagonal
// (MatrixGenerator.cpp:143-153). The kernel is write-bound (every stored
// element written once) plus one FP64 exp per element.
#include <algorithm>

#include "hs_internal.h"

namespace {

constexpr int RBF_THREADS = 256;

__device__ inline void decode_block_col(i64 id, int n, int* I, int* J) {
    const double nn = 2.0 * n + 1.0;
    i64 j = (i64)((nn - sqrt(nn * nn - 8.0 * (double)id)) * 0.5);
    if (j < 0) j = 0;
    while (j > 0 && j * n - j * (j - 1) / 2 > id) --j;
    while ((j + 1) * n - (j + 1) * j / 2 <= id) ++j;
    *J = (int)j;
    *I = (int)(j + (id - (j * n - j * (j - 1) / 2)));
}

// Fast path: b in {64, 128, 256}, feature count D fixed at compile time.
// The block row's features (b x D, contiguous in X) are staged in LDS once;
// thread t owns column j = t % b (its features in registers) and rows
// i = t / b + s * (256 / b), and all lanes of a wave share the row, so x_i
// comes from LDS as a broadcast. Each row of the block is one coalesced run
// of non-temporal stores (the matrix is written once here and read later).
template <int BS, int D>
__global__ __launch_bounds__(RBF_THREADS) void rbf_fast_kernel(const double* __restrict__ X, i64 N, int n,
                                                                double scale, double sf2, double sn2,
                                                                double* __restrict__ A) {
    constexpr int RP = RBF_THREADS / BS;  // row phases
    __shared__ double xs[BS * D];
    const i64 id = blockIdx.x;
    int I, J;
    decode_block_col(id, n, &I, &J);
    double* blk = A + id * (i64)BS * BS;
    const i64 rowBase = (i64)I * BS;
    const i64 rowValid = std::min<i64>((i64)BS, N - rowBase);
    for (int e = threadIdx.x; e < BS * D; e += RBF_THREADS) xs[e] = (e / D < rowValid) ? X[rowBase * D + e] : 0.0;
    const int j = threadIdx.x % BS;
    const int ph = threadIdx.x / BS;
    const i64 jg = (i64)J * BS + j;
    const bool jvalid = jg < N;
    double xj[D];
#pragma unroll
    for (int k = 0; k < D; ++k) xj[k] = jvalid ? X[jg * D + k] : 0.0;
    __syncthreads();
    for (int i = ph; i < BS; i += RP) {
        const i64 ig = rowBase + i;
        double v = 0.0;
        if (i < rowValid && jvalid) {
            const double* xi = xs + i * D;
            double dist = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double tmp = xi[k] - xj[k];
                dist += tmp * tmp;
            }
            v = sf2 * exp(scale * dist);
            if (ig == jg) v += sn2;
        }
        __builtin_nontemporal_store(v, blk + (i64)i * BS + j);
    }
}

// Any b and d: element loop, features read from global (cached).
__global__ __launch_bounds__(RBF_THREADS) void rbf_generic_kernel(const double* __restrict__ X, i64 N, int d, int b,
                                                                   int n, double scale, double sf2, double sn2,
                                                                   double* __restrict__ A) {
    const i64 id = blockIdx.x;
    int I, J;
    decode_block_col(id, n, &I, &J);
    double* blk = A + id * (i64)b * b;
    for (int e = threadIdx.x; e < b * b; e += RBF_THREADS) {
        const int i = e / b, j = e % b;
        const i64 ig = (i64)I * b + i, jg = (i64)J * b + j;
        double v = 0.0;
        if (ig < N && jg < N) {
            double dist = 0.0;
            for (int k = 0; k < d; ++k) {
                const double tmp = X[ig * d + k] - X[jg * d + k];
                dist += tmp * tmp;
            }
            v = sf2 * exp(scale * dist);
            if (ig == jg) v += sn2;
        }
        blk[e] = v;
    }
}

// 8-lag window features of a series (MatrixGenerator.cpp:77-99):
// X[i][k] = t[i + k - 7], zero for negative indices.
__global__ void window_features_kernel(const double* __restrict__ t, i64 N, double* __restrict__ X) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * 8) return;
    const i64 i = e / 8;
    const int k = (int)(e % 8);
    const i64 src = i + k - 7;
    X[e] = src >= 0 ? t[src] : 0.0;
}

}  // namespace

int hs_launch_rbf(hs_handle_t h, const double* X, i64 N, int d, int b, double ell, double sf2, double sn2, double* A) {
    const int n = (int)((N + b - 1) / b);
    const i64 blocks = (i64)n * (n + 1) / 2;
    if (blocks > 0x7fffffffLL) {
        hs_set_error("RBF: too many blocks");
        return HS_ERR_UNSUPPORTED;
    }
    const double scale = -0.5 / (ell * ell);
    const dim3 grid((unsigned)blocks), blk(RBF_THREADS);
#define HS_RBF_CASE(BS, D)                                                                         \
    if (b == BS && d == D) {                                                                       \
        rbf_fast_kernel<BS, D><<<grid, blk, 0, h->stream>>>(X, N, n, scale, sf2, sn2, A);          \
        HS_LAUNCH_CHECK();                                                                         \
        return HS_OK;                                                                              \
    }
    HS_RBF_CASE(64, 8)
    HS_RBF_CASE(128, 8)
    HS_RBF_CASE(256, 8)
    HS_RBF_CASE(64, 16)
    HS_RBF_CASE(128, 16)
    HS_RBF_CASE(256, 16)
#undef HS_RBF_CASE
    rbf_generic_kernel<<<grid, blk, 0, h->stream>>>(X, N, d, b, n, scale, sf2, sn2, A);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

extern "C" int hs_rbf_assemble_f64(hs_handle_t h, const double* X, int64_t N, int d, int b, double ell, double sf2,
                                   double sn2, double* A) {
    if (!h) return -1;
    if (!X) return -2;
    if (N <= 0) return -3;
    if (d <= 0) return -4;
    if (b <= 0) return -5;
    if (!(ell > 0.0)) return -6;
    if (!A) return -9;
    return hs_launch_rbf(h, X, N, d, b, ell, sf2, sn2, A);
}

size_t hs_ws_rbf_window(i64 N) { return sizeof(double) * (size_t)N * 8; }

extern "C" int hs_rbf_window_f64(hs_handle_t h, const double* t, int64_t N, int b, double* A) {
    if (!h) return -1;
    if (!t) return -2;
    if (N <= 0) return -3;
    if (b <= 0) return -4;
    if (!A) return -5;
    double* X = nullptr;
    int rc = hs_work(h, hs_ws_rbf_window(N), (void**)&X);
    if (rc) return rc;
    window_features_kernel<<<dim3((unsigned)((N * 8 + 255) / 256)), dim3(256), 0, h->stream>>>(t, N, X);
    HS_LAUNCH_CHECK();
    return hs_launch_rbf(h, X, N, 8, b, 1.0, 1.0, 0.01, A);
}
