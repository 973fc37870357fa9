// Gaussian-process prediction after the solve: the test/training kernel
// matrix K* and the predictive mean K* alpha.
//
// The reference (GaussianProcess::start, GaussianProcess.cpp:42-59) fills a
// host K* of N_test x N entries on the CPU queue
// (MatrixGenerator::generateTestKernelMatrix, MatrixGenerator.cpp:163-213:
//   K*[j N + i] = vl exp(-0.5 / l^2 * sum_k (xtrain_i,k - xtest_j,k)^2),
// distance summed in feature order) and then runs a one-work-item-per-row
// mat-vec over it (MatrixVectorOperations::matrixVectorGP,
// MatrixVectorOperations.cpp:379-395). Both steps have device entries here
// with the same layouts; hs_gp_predict_f64 fuses them, so K* never reaches
// HBM: the prediction is then bound by the FP64 exp of each (test, train)
// pair instead of by 16 bytes of K* traffic per pair.
#include <algorithm>

#include "hs_internal.h"

namespace {

constexpr int GT = 256;      // threads per work-group
constexpr int RT = 16;       // test rows per work-group of the K* assembly
constexpr int DMAX = 64;     // features held in LDS per test point

// squared distance between training row xi (registers or global) and test
// row xs (LDS), in feature order
template <int D>
__device__ inline double dist2(const double* xi, const double* xs, int d) {
    double dist = 0.0;
    if (D > 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double tmp = xi[k] - xs[k];
            dist += tmp * tmp;
        }
    } else {
        for (int k = 0; k < d; ++k) {
            const double tmp = xi[k] - xs[k];
            dist += tmp * tmp;
        }
    }
    return dist;
}

// K* assembly: thread owns training column i (its features in registers for
// D > 0), the work-group's RT test rows sit in LDS; every row j of the tile
// is written as one coalesced run of columns, with non-temporal stores (K*
// is written once and read once by the mat-vec).
template <int D>
__global__ __launch_bounds__(GT) void rbf_cross_kernel(const double* __restrict__ Xtr, i64 N,
                                                       const double* __restrict__ Xte, i64 M, int d, double scale,
                                                       double sf2, double* __restrict__ K) {
    __shared__ double xs[RT * DMAX];
    const int dd = D > 0 ? D : d;
    const i64 i = (i64)blockIdx.x * GT + threadIdx.x;
    const i64 j0 = (i64)blockIdx.y * RT;
    const int rows = (int)std::min<i64>(RT, M - j0);
    for (int e = threadIdx.x; e < rows * dd; e += GT) xs[e] = Xte[j0 * dd + e];
    __syncthreads();
    if (i >= N) return;
    double xr[D > 0 ? D : 1];
    const double* xi = Xtr + i * dd;
    if (D > 0) {
#pragma unroll
        for (int k = 0; k < (D > 0 ? D : 1); ++k) xr[k] = xi[k];
    }
    for (int r = 0; r < rows; ++r) {
        const double dist = dist2<D>(D > 0 ? xr : xi, xs + r * dd, dd);
        __builtin_nontemporal_store(sf2 * exp(scale * dist), K + (j0 + r) * N + i);
    }
}

// y_r = sum_k A[r n + k] x_k for m rows of length n (row-major, any n): one
// wave per row, 8-byte loads (rows of odd length start 8-byte aligned), eight
// loads in flight per lane, a wave reduction at the end.
__global__ __launch_bounds__(GT) void gemv_gp_kernel(const double* __restrict__ A, const double* __restrict__ x,
                                                     double* __restrict__ y, i64 n, i64 m) {
    const i64 r = (i64)blockIdx.x * (GT / 64) + (threadIdx.x >> 6);
    if (r >= m) return;  // uniform per wave
    const int lane = threadIdx.x & 63;
    const double* row = A + r * n;
    double acc = 0.0;
    i64 k = lane;
    for (; k + 7 * 64 < n; k += 8 * 64) {
        double a[8], v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a[u] = __builtin_nontemporal_load(row + k + u * 64);
            v[u] = x[k + u * 64];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += a[u] * v[u];
    }
    for (; k < n; k += 64) acc += row[k] * x[k];
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) acc += __shfl_xor(acc, s, 64);
    if (lane == 0) y[r] = acc;
}

// Fused prediction, pass 1: work-group (test tile t, training split s) owns
// GT test points, one per thread (features in registers), and walks its
// training range in LDS tiles of TP points (features and sf2 alpha_i), every
// thread reading the same tile entry at the same time (LDS broadcast). Thread
// j adds sf2 alpha_i exp(scale ||x_i - x*_j||^2) over the range in ascending
// i and writes part[j * S + s].
constexpr int TP = 64;

template <int D>
__global__ __launch_bounds__(GT) void gp_predict_partial_kernel(const double* __restrict__ Xtr, i64 N,
                                                                const double* __restrict__ Xte, i64 M, int d,
                                                                double scale, double sf2,
                                                                const double* __restrict__ alpha, i64 chunk,
                                                                double* __restrict__ part) {
    __shared__ double xt[TP * DMAX];
    __shared__ double at[TP];
    const int dd = D > 0 ? D : d;
    const int S = gridDim.y;
    const i64 j = (i64)blockIdx.x * GT + threadIdx.x;
    const bool valid = j < M;
    double xr[D > 0 ? D : 1];
    if (D > 0) {
#pragma unroll
        for (int k = 0; k < (D > 0 ? D : 1); ++k) xr[k] = valid ? Xte[j * dd + k] : 0.0;
    }
    const double* xj = Xte + (valid ? j : 0) * dd;
    const i64 i0 = (i64)blockIdx.y * chunk, i1 = std::min<i64>(N, i0 + chunk);
    double acc = 0.0;
    for (i64 t0 = i0; t0 < i1; t0 += TP) {
        const int cnt = (int)std::min<i64>(TP, i1 - t0);
        __syncthreads();  // the previous tile has been read by every thread
        for (int e = threadIdx.x; e < cnt * dd; e += GT) xt[e] = Xtr[t0 * dd + e];
        if ((int)threadIdx.x < cnt) at[threadIdx.x] = sf2 * alpha[t0 + threadIdx.x];
        __syncthreads();
#pragma unroll 2
        for (int p = 0; p < cnt; ++p) {
            double dist;
            if (D > 0) {
                dist = dist2<D>(xt + p * dd, xr, dd);
            } else {
                dist = dist2<0>(xt + p * dd, xj, dd);
            }
            acc += at[p] * exp(scale * dist);
        }
    }
    if (valid) part[j * S + blockIdx.y] = acc;
}

// pass 2: mean_j = sum_s part[j S + s], s ascending
__global__ __launch_bounds__(GT) void gp_predict_sum_kernel(const double* __restrict__ part, int S, i64 M,
                                                            double* __restrict__ mean) {
    const i64 j = (i64)blockIdx.x * GT + threadIdx.x;
    if (j >= M) return;
    double v = 0.0;
    for (int s = 0; s < S; ++s) v += part[j * S + s];
    mean[j] = v;
}

// 8-lag window features of a series (MatrixGenerator.cpp:77-99, 257-272):
// X[i][k] = t[i + k - 7], zero for negative indices.
__global__ void window8_kernel(const double* __restrict__ t, i64 N, double* __restrict__ X) {
    const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * 8) return;
    const i64 i = e / 8;
    const int k = (int)(e % 8);
    const i64 src = i + k - 7;
    X[e] = src >= 0 ? t[src] : 0.0;
}

int check_cross(hs_handle_t h, const double* Xtr, int64_t N, const double* Xte, int64_t M, int d, double ell) {
    if (!h) return -1;
    if (!Xtr) return -2;
    if (N <= 0) return -3;
    if (!Xte) return -4;
    if (M <= 0) return -5;
    if (d <= 0 || d > DMAX) {
        hs_set_error("GP prediction: feature count must be in 1..64");
        return -6;
    }
    if (!(ell > 0.0)) return -7;
    return HS_OK;
}

}  // namespace

extern "C" int hs_rbf_cross_f64(hs_handle_t h, const double* Xtrain, int64_t N, const double* Xtest, int64_t M, int d,
                                double ell, double sf2, double* Kstar) {
    int rc = check_cross(h, Xtrain, N, Xtest, M, d, ell);
    if (rc) return rc;
    if (!Kstar) return -9;
    const i64 gy = (M + RT - 1) / RT, gx = (N + GT - 1) / GT;
    if (gy > 65535 || gx > 0x7fffffffLL) {
        hs_set_error("hs_rbf_cross_f64: too many test points for one launch");
        return HS_ERR_UNSUPPORTED;
    }
    const double scale = -0.5 / (ell * ell);
    const dim3 grid((unsigned)gx, (unsigned)gy), blk(GT);
    switch (d) {
        case 8: rbf_cross_kernel<8><<<grid, blk, 0, h->stream>>>(Xtrain, N, Xtest, M, d, scale, sf2, Kstar); break;
        case 16: rbf_cross_kernel<16><<<grid, blk, 0, h->stream>>>(Xtrain, N, Xtest, M, d, scale, sf2, Kstar); break;
        default: rbf_cross_kernel<0><<<grid, blk, 0, h->stream>>>(Xtrain, N, Xtest, M, d, scale, sf2, Kstar);
    }
    HS_LAUNCH_CHECK();
    return HS_OK;
}

extern "C" int hs_gemv_rows_f64(hs_handle_t h, const double* A, const double* x, double* y, int64_t n, int64_t m) {
    if (!h) return -1;
    if (!A) return -2;
    if (!x) return -3;
    if (!y) return -4;
    if (n <= 0) return -5;
    if (m <= 0) return -6;
    const i64 grid = (m + GT / 64 - 1) / (GT / 64);
    if (grid > 0x7fffffffLL) return -6;
    gemv_gp_kernel<<<dim3((unsigned)grid), dim3(GT), 0, h->stream>>>(A, x, y, n, m);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

// training splits: about 2048 work-groups in all, at least TP training
// points each, at most 65535; returns their count and size (*chunk)
static int predict_splits(i64 N, i64 M, i64* chunk) {
    const i64 tiles = (M + GT - 1) / GT;
    const i64 maxS = std::max<i64>(1, (N + TP - 1) / TP);
    const i64 S = std::min<i64>(std::min<i64>(maxS, 65535), std::max<i64>(1, (2048 + tiles - 1) / tiles));
    *chunk = (N + S - 1) / S;
    return (int)((N + *chunk - 1) / *chunk);
}

size_t hs_ws_gp_predict(i64 N, i64 M) {
    i64 chunk = 0;
    return sizeof(double) * (size_t)(M * predict_splits(N, M, &chunk));
}

extern "C" int hs_gp_predict_f64(hs_handle_t h, const double* Xtrain, int64_t N, const double* Xtest, int64_t M, int d,
                                 double ell, double sf2, const double* alpha, double* mean) {
    int rc = check_cross(h, Xtrain, N, Xtest, M, d, ell);
    if (rc) return rc;
    if (!alpha) return -9;
    if (!mean) return -10;
    const i64 tiles = (M + GT - 1) / GT;
    if (tiles > 0x7fffffffLL) return -5;
    i64 chunk = 0;
    const int Sn = predict_splits(N, M, &chunk);
    double* part = nullptr;
    if ((rc = hs_work(h, sizeof(double) * (size_t)(M * Sn), (void**)&part))) return rc;
    const double scale = -0.5 / (ell * ell);
    const dim3 grid((unsigned)tiles, (unsigned)Sn), blk(GT);
    switch (d) {
        case 8:
            gp_predict_partial_kernel<8><<<grid, blk, 0, h->stream>>>(Xtrain, N, Xtest, M, d, scale, sf2, alpha, chunk,
                                                                      part);
            break;
        case 16:
            gp_predict_partial_kernel<16><<<grid, blk, 0, h->stream>>>(Xtrain, N, Xtest, M, d, scale, sf2, alpha,
                                                                       chunk, part);
            break;
        default:
            gp_predict_partial_kernel<0><<<grid, blk, 0, h->stream>>>(Xtrain, N, Xtest, M, d, scale, sf2, alpha, chunk,
                                                                      part);
    }
    HS_LAUNCH_CHECK();
    gp_predict_sum_kernel<<<dim3((unsigned)((M + GT - 1) / GT)), blk, 0, h->stream>>>(part, Sn, M, mean);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

extern "C" int hs_window_features_f64(hs_handle_t h, const double* t, int64_t N, double* X) {
    if (!h) return -1;
    if (!t) return -2;
    if (N <= 0) return -3;
    if (!X) return -4;
    const i64 grid = (N * 8 + GT - 1) / GT;
    if (grid > 0x7fffffffLL) return -3;
    window8_kernel<<<dim3((unsigned)grid), dim3(GT), 0, h->stream>>>(t, N, X);
    HS_LAUNCH_CHECK();
    return HS_OK;
}
