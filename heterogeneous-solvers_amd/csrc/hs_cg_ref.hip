// CG in the reference's own floating-point order (share 1.0, GPU queue).
//
// The reference's GPU CG (CG.cpp:38-163) fixes the order of every operation
// through its kernels; this file performs the same operations in the same
// order, so that a run reproduces its iterates bit for bit (up to the one
// thing the source leaves open, whether `a += b * c` is contracted into a
// fused multiply-add, which every entry here takes as an argument):
//   * mat-vec, matrixVectorBlock_GPU (MatrixVectorOperations.cpp:95-192):
//     row i is summed from 0 over block columns J = 0..n-1 and columns
//     j = 0..b-1 in order; stored block (I,J) read row-wise for J <= I, block
//     (J,I) read transposed for J > I.
//   * dot product, scalarProduct + sumFinalScalarProduct
//     (VectorOperations.cpp:84-185): wg items per work-group over 2 wg
//     elements, item l = x_l y_l (+) x_{l+wg} y_{l+wg}, a halving tree
//     cache[l] += cache[l+s]; then the same over the partials with wgf items.
//   * updates, scaleAndAddVectorBlock (x + y a) and subVectorBlock (x - y).
// This is the parity mode. The fast CG (hs_cg.hip) reads each stored block
// once and sums in its own fixed order.
//
// Layout of the mat-vec: one work-group per band of up to 128 rows of a
// block row I (two bands per block row at b = 256, so that every CU gets
// one), thread t owns one row and keeps its running sum in a register for
// the whole pass. The pass walks the block row as a stream of CH-column
// chunks. Row-wise blocks are loaded coalesced (CH consecutive doubles of
// each row per 16 lanes), transposed into a double-buffered LDS tile and
// read back as column t; transposed blocks are read directly, lane t taking
// element t of each stored row. The next chunk's loads are issued before
// the current chunk's multiply-adds. When b is a multiple of the chunk and
// of the band, the loads carry no bounds tests (FULL): guarded, each load
// became its own branch and the loads of a chunk no longer overlapped
// (2.5x slower at b = 128).
#include <algorithm>
#include <climits>

#include "hs_internal.h"

namespace {

// a + b * c in the requested rounding: contracted (one rounding) or not
// (two). The pragma keeps the compiler from fusing the second form.
template <bool FMA>
__device__ __forceinline__ double madd(double acc, double a, double x) {
    if constexpr (FMA) {
        return __builtin_fma(a, x, acc);
    } else {
#pragma clang fp contract(off)
        const double p = a * x;
        return acc + p;
    }
}

__device__ __forceinline__ double add_nc(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}

template <int CH, bool FMA, bool FULL>
__global__ __launch_bounds__(256) void symv_ordered_kernel(const double* __restrict__ A, const double* __restrict__ x,
                                                           double* __restrict__ y, int n, int b, int RB,
                                                           const i64* __restrict__ stop, i64 it) {
    if (stop && it > *stop) return;
    extern __shared__ double tile[];  // [2][CH][RB + 1]: the odd pitch spreads a chunk row over the banks
    const int bands = (b + RB - 1) / RB;
    const int I = blockIdx.x / bands;
    const int rb0 = (blockIdx.x % bands) * RB;  // the band's first row in the block row
    const int rows = min(RB, b - rb0);
    const int t = threadIdx.x;
    const int T = blockDim.x;
    const int cpb = (b + CH - 1) / CH;  // chunks per block
    const int K = n * cpb;
    const i64 bb = (i64)b * b;
    const int P = RB + 1;
    double R[CH];
    double acc = 0.0;

    // loads of chunk k into R: row-wise chunks as a rows x CH sub-matrix in
    // row-major element order e = r * CH + c (lane e % T), transposed chunks
    // as rows c0..c0+CH-1 of the stored block, lane t taking column rb0 + t
    auto load = [&](int k) {
        const int J = k / cpb, c0 = (k % cpb) * CH;
        if (J <= I) {
            const double* blk = A + hs_block_id(I, J, n) * bb + (i64)rb0 * b;
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = t + u * T;
                const int r = e / CH, c = e % CH;
                if constexpr (FULL) R[u] = __builtin_nontemporal_load(blk + (i64)r * b + c0 + c);
                else R[u] = (r < rows && c0 + c < b) ? __builtin_nontemporal_load(blk + (i64)r * b + c0 + c) : 0.0;
            }
        } else {
            const double* blk = A + hs_block_id(J, I, n) * bb + rb0;
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                if constexpr (FULL) R[u] = __builtin_nontemporal_load(blk + (i64)(c0 + u) * b + t);
                else R[u] = (t < rows && c0 + u < b) ? __builtin_nontemporal_load(blk + (i64)(c0 + u) * b + t) : 0.0;
            }
        }
    };

    load(0);
    for (int k = 0; k < K; ++k) {
        const int J = k / cpb, c0 = (k % cpb) * CH;
        const int cw = min(CH, b - c0);
        double* tl = tile + (k & 1) * CH * P;
        const bool rowwise = J <= I;
        if (rowwise) {
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = t + u * T;
                const int r = e / CH, c = e % CH;
                if (FULL || r < rows) tl[c * P + r] = R[u];
            }
        }
        __syncthreads();
        double V[CH];
        if (rowwise) {
#pragma unroll
            for (int c = 0; c < CH; ++c) V[c] = (FULL || t < rows) ? tl[c * P + t] : 0.0;
        } else {
#pragma unroll
            for (int c = 0; c < CH; ++c) V[c] = R[c];
        }
        if (k + 1 < K) load(k + 1);
        const double* xs = x + (i64)J * b + c0;
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (FULL || c < cw) acc = madd<FMA>(acc, V[c], xs[c]);
    }
    if (FULL || t < rows) y[(i64)I * b + rb0 + t] = acc;
}

// scalarProduct: one work-group of wg items per 2 wg elements of x y over
// [off, off + len); the partial of work-group g goes to part[g].
template <bool FMA>
__global__ void dot_tree_kernel(const double* __restrict__ x, const double* __restrict__ y, i64 off, i64 len,
                                double* __restrict__ part, const i64* __restrict__ stop, i64 it) {
    if (stop && it > *stop) return;
    extern __shared__ double cache[];
    const unsigned wg = blockDim.x, l = threadIdx.x;
    const i64 gi = off + (i64)blockIdx.x * (wg * 2) + l;
    double c = 0.0;
    if (gi < off + len) c = madd<FMA>(c, x[gi], y[gi]);
    if (gi + wg < off + len) c = madd<FMA>(c, x[gi + wg], y[gi + wg]);
    cache[l] = c;
    __syncthreads();
    for (unsigned s = wg / 2; s > 0; s >>= 1) {
        if (l < s) cache[l] = add_nc(cache[l], cache[l + s]);
        __syncthreads();
    }
    if (l == 0) part[blockIdx.x] = cache[0];
}

// sumFinalScalarProduct: wgf items over G <= 2 wgf partials; the sum goes to
// *out.
__global__ void dot_final_kernel(const double* __restrict__ part, int G, double* __restrict__ out,
                                 const i64* __restrict__ stop, i64 it) {
    if (stop && it > *stop) return;
    extern __shared__ double cache[];
    const unsigned wgf = blockDim.x, l = threadIdx.x;
    double c = 0.0;
    if ((int)l < G) c = add_nc(c, part[l]);
    if ((int)(l + wgf) < G) c = add_nc(c, part[l + wgf]);
    cache[l] = c;
    __syncthreads();
    for (unsigned s = wgf / 2; s > 0; s >>= 1) {
        if (l < s) cache[l] = add_nc(cache[l], cache[l + s]);
        __syncthreads();
    }
    if (l == 0) *out = cache[0];
}

struct RefState {
    double delta[2];  // delta_new of even / odd iteration parity
    double dq;        // d^T q of the current iteration
    double delta0;
    double eps2;
    double pad[3];
    i64 stop;  // last executed iteration; LLONG_MAX while running
};

constexpr int VT = 256;  // vector-kernel work-group size

__global__ __launch_bounds__(VT) void ref_zero_kernel(double* __restrict__ v, i64 len) {
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i < len) v[i] = 0.0;
}

// r = rhs - r (subVectorBlock); with init, also d = r (initCG's copy)
__global__ __launch_bounds__(VT) void ref_sub_kernel(const double* __restrict__ rhs, double* __restrict__ r,
                                                     double* __restrict__ d, i64 len, const i64* __restrict__ stop,
                                                     i64 it) {
    if (stop && it > *stop) return;
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i >= len) return;
    const double v = rhs[i] - r[i];
    r[i] = v;
    if (d) d[i] = v;
}

__global__ void ref_init_state_kernel(RefState* st, double eps2) {
    const double s = st->delta[0];
    st->delta[1] = s;
    st->delta0 = s;
    st->eps2 = eps2;
    st->dq = 0.0;
    st->stop = (s > eps2 * s) ? LLONG_MAX : -1;
}

// x = x + d alpha, alpha = delta_new / d^T q (compute_alpha, update_x)
template <bool FMA>
__global__ __launch_bounds__(VT) void ref_update_x_kernel(RefState* st, const double* __restrict__ d,
                                                          double* __restrict__ x, i64 len, i64 it) {
    if (it > st->stop) return;
    const double alpha = st->delta[it & 1] / st->dq;
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i < len) x[i] = madd<FMA>(x[i], d[i], alpha);
}

// r = r + q (-1.0 alpha) (update_r)
template <bool FMA>
__global__ __launch_bounds__(VT) void ref_update_r_kernel(RefState* st, const double* __restrict__ q,
                                                          double* __restrict__ r, i64 len, i64 it) {
    if (it > st->stop) return;
    const double alpha = st->delta[it & 1] / st->dq;
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i < len) r[i] = madd<FMA>(r[i], q[i], -1.0 * alpha);
}

// d = r + d beta, beta = delta_new / delta_old (compute_d); the stop test of
// the loop head (CG.cpp:78) for the next iteration
template <bool FMA>
__global__ __launch_bounds__(VT) void ref_update_d_kernel(RefState* st, const double* __restrict__ r,
                                                          double* __restrict__ d, i64 len, i64 it) {
    if (it > st->stop) return;
    const double dn = st->delta[(it + 1) & 1];
    const double beta = dn / st->delta[it & 1];
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i < len) d[i] = madd<FMA>(r[i], d[i], beta);
    if (blockIdx.x == 0 && threadIdx.x == 0 && !(dn > st->eps2 * st->delta0)) st->stop = it;
}

bool pow2_in(int v, int hi) { return v >= 1 && v <= hi && (v & (v - 1)) == 0; }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool hs_symv_ordered_ok(int b) { return b >= 1 && b <= 256; }

template <int CH, bool FMA, bool FULL>
static int launch_ordered(hs_handle_t h, const double* A, const double* x, double* y, int n, int b, const i64* stop,
                          i64 it) {
    const int RB = std::min(b, 128);             // rows per work-group
    const int bands = (b + RB - 1) / RB;
    // two CH x (RB + 1) tiles: 66 KB at CH = 32
    const size_t lds = sizeof(double) * 2 * CH * (RB + 1);
    // set per launch: the attribute belongs to the device current at the call
    HS_HIP(hipFuncSetAttribute((const void*)symv_ordered_kernel<CH, FMA, FULL>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)(sizeof(double) * 2 * CH * 129)));
    const int T = std::max(64, (RB + 63) / 64 * 64);
    symv_ordered_kernel<CH, FMA, FULL><<<dim3((unsigned)(n * bands)), dim3(T), lds, h->stream>>>(A, x, y, n, b, RB,
                                                                                              stop, it);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

int hs_launch_symv_ordered(hs_handle_t h, const double* A, const double* x, double* y, int n, int b, int fma,
                           const i64* stop, i64 it) {
    // FULL: every band has 64k rows and every chunk 32 columns
    const bool full = b % 32 == 0 && (b <= 128 ? b % 64 == 0 : b % 128 == 0);
    if (full)
        return fma ? launch_ordered<32, true, true>(h, A, x, y, n, b, stop, it)
                   : launch_ordered<32, false, true>(h, A, x, y, n, b, stop, it);
    return fma ? launch_ordered<32, true, false>(h, A, x, y, n, b, stop, it)
               : launch_ordered<32, false, false>(h, A, x, y, n, b, stop, it);
}

// partial count of scalarProduct over len elements with wg items
static i64 dot_groups(i64 len, int wg) { return (len / 2 + wg - 1) / wg; }

static int launch_dot_tree(hs_handle_t h, const double* x, const double* y, i64 off, i64 len, int wg, int wgf,
                           int fma, double* part, double* out, const i64* stop, i64 it) {
    const i64 G = dot_groups(len, wg);
    if (G > 2 * (i64)wgf) {
        hs_set_error("Workgroup size for final scalar product is too small");
        return HS_ERR_UNSUPPORTED;
    }
    if (G > 0) {
        if (fma)
            dot_tree_kernel<true><<<dim3((unsigned)G), dim3(wg), sizeof(double) * wg, h->stream>>>(x, y, off, len,
                                                                                                 part, stop, it);
        else
            dot_tree_kernel<false><<<dim3((unsigned)G), dim3(wg), sizeof(double) * wg, h->stream>>>(x, y, off, len,
                                                                                                  part, stop, it);
        HS_LAUNCH_CHECK();
    }
    dot_final_kernel<<<1, dim3(wgf), sizeof(double) * wgf, h->stream>>>(part, (int)G, out, stop, it);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

size_t hs_ws_dot_ref(i64 len, int wg) { return sizeof(double) * (size_t)(dot_groups(len, wg) + 1); }

size_t hs_ws_cg_ref(i64 N, int b, int wg) {
    const i64 rows = (N + b - 1) / b * b;
    return sizeof(double) * (size_t)(3 * rows + dot_groups(rows, wg) + 8) + sizeof(RefState) + 256;
}

extern "C" int hs_symv_ref_order_f64(hs_handle_t h, const double* A, const double* x, double* y, int64_t N, int b,
                                     int fma) {
    if (!h) return -1;
    if (!A) return -2;
    if (!x) return -3;
    if (!y) return -4;
    if (N <= 0) return -5;
    if (!hs_symv_ordered_ok(b)) {
        hs_set_error("hs_symv_ref_order_f64: block size must be 1..256");
        return -6;
    }
    if (fma != 0 && fma != 1) return -7;
    const i64 n = (N + b - 1) / b;
    if (n > INT_MAX) return -5;
    return hs_launch_symv_ordered(h, A, x, y, (int)n, b, fma, nullptr, 0);
}

extern "C" int hs_dot_ref_order_f64(hs_handle_t h, const double* x, const double* y, int b, int blockStart,
                                    int blockCount, int wg, int wgf, int fma, double* result_dev) {
    if (!h) return -1;
    if (!x) return -2;
    if (!y) return -3;
    if (b <= 0) return -4;
    if (blockStart < 0) return -5;
    if (blockCount < 0) return -6;
    if (!pow2_in(wg, 1024)) return -7;
    if (!pow2_in(wgf, 1024)) return -8;
    if (fma != 0 && fma != 1) return -9;
    if (!result_dev) return -10;
    const i64 len = (i64)blockCount * b;
    double* part = nullptr;
    int rc = hs_work(h, hs_ws_dot_ref(len, wg), (void**)&part);
    if (rc) return rc;
    return launch_dot_tree(h, x, y, (i64)blockStart * b, len, wg, wgf, fma, part, result_dev, nullptr, 0);
}

static int cg_ref_run(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b, double eps,
                      int64_t imax, int wg, int wgf, int fma, int64_t* iterations, double* delta_final,
                      double* delta_zero, double* trace) {
    if (!h) return -1;
    if (!A) return -2;
    if (!rhs) return -3;
    if (!x) return -4;
    if (N <= 0) return -5;
    if (!hs_symv_ordered_ok(b)) {
        hs_set_error("hs_cg_ref_order_f64: block size must be 1..256");
        return -6;
    }
    if (!(eps >= 0.0)) return -7;
    if (imax < 0) return -8;
    if (!pow2_in(wg, 1024)) return -9;
    if (!pow2_in(wgf, 1024)) return -10;
    if (fma != 0 && fma != 1) return -11;
    const i64 n = (N + b - 1) / b;
    if (n > INT_MAX) return -5;
    const i64 rows = n * b;
    if (dot_groups(rows, wg) > 2 * (i64)wgf) {
        hs_set_error("Workgroup size for final scalar product is too small");
        return HS_ERR_UNSUPPORTED;
    }
    double* w = nullptr;
    int rc = hs_work(h, hs_ws_cg_ref(N, b, wg), (void**)&w);
    if (rc) return rc;
    double* d = w;
    double* q = w + rows;
    double* r = w + 2 * rows;
    double* part = w + 3 * rows;
    RefState* st = (RefState*)(((uintptr_t)(part + dot_groups(rows, wg) + 8) + 63) & ~(uintptr_t)63);
    const i64* stop = &st->stop;
    const dim3 vg((unsigned)((rows + VT - 1) / VT)), vb(VT);

    // initCG (CG.cpp:277-340): x = 0, r = b - A x, d = r, delta_0 = r.r
    ref_zero_kernel<<<vg, vb, 0, h->stream>>>(x, rows);
    HS_LAUNCH_CHECK();
    if ((rc = hs_launch_symv_ordered(h, A, x, r, (int)n, b, fma, nullptr, 0))) return rc;
    ref_sub_kernel<<<vg, vb, 0, h->stream>>>(rhs, r, d, rows, nullptr, 0);
    HS_LAUNCH_CHECK();
    if ((rc = launch_dot_tree(h, r, r, 0, rows, wg, wgf, fma, part, &st->delta[0], nullptr, 0))) return rc;
    ref_init_state_kernel<<<1, 1, 0, h->stream>>>(st, eps * eps);
    HS_LAUNCH_CHECK();

    const int CHUNK = 8;
    volatile i64* pinnedStop = (volatile i64*)h->pinned;
    i64 enq = 0;
    hipEvent_t evs[2] = {h->events[0], h->events[1]};
    IterTrace tr;
    if ((rc = tr.start(trace, imax, CHUNK))) return rc;
    i64 chunkFirst[2] = {0, 0}, chunkEnd[2] = {0, 0};
    auto enqueue_chunk = [&]() -> int {
        const i64 end = std::min<i64>(enq + CHUNK, imax);
        for (; enq < end; ++enq) {
            const i64 it = enq;
            int e;
            // q = A d ; alpha = delta_new / d.q ; x = x + alpha d
            if ((e = tr.mark(h->stream, it, 0))) return e;
            if ((e = hs_launch_symv_ordered(h, A, d, q, (int)n, b, fma, stop, it))) return e;
            if ((e = tr.mark(h->stream, it, 1))) return e;
            if ((e = launch_dot_tree(h, d, q, 0, rows, wg, wgf, fma, part, &st->dq, stop, it))) return e;
            if (fma)
                ref_update_x_kernel<true><<<vg, vb, 0, h->stream>>>(st, d, x, rows, it);
            else
                ref_update_x_kernel<false><<<vg, vb, 0, h->stream>>>(st, d, x, rows, it);
            HS_LAUNCH_CHECK();
            if ((e = tr.mark(h->stream, it, 2))) return e;
            if (it % 50 == 0) {
                // true residual (CG.cpp:95-101, 444-473)
                if ((e = hs_launch_symv_ordered(h, A, x, r, (int)n, b, fma, stop, it))) return e;
                ref_sub_kernel<<<vg, vb, 0, h->stream>>>(rhs, r, nullptr, rows, stop, it);
            } else if (fma) {
                ref_update_r_kernel<true><<<vg, vb, 0, h->stream>>>(st, q, r, rows, it);
            } else {
                ref_update_r_kernel<false><<<vg, vb, 0, h->stream>>>(st, q, r, rows, it);
            }
            HS_LAUNCH_CHECK();
            if ((e = tr.mark(h->stream, it, 3))) return e;
            // delta_new = r.r ; beta ; d = r + beta d ; stop test
            if ((e = launch_dot_tree(h, r, r, 0, rows, wg, wgf, fma, part, &st->delta[(it + 1) & 1], stop, it)))
                return e;
            if (fma)
                ref_update_d_kernel<true><<<vg, vb, 0, h->stream>>>(st, r, d, rows, it);
            else
                ref_update_d_kernel<false><<<vg, vb, 0, h->stream>>>(st, r, d, rows, it);
            HS_LAUNCH_CHECK();
            if ((e = tr.mark(h->stream, it, 4))) return e;
        }
        return HS_OK;
    };
    auto read_slot = [&](int s, i64 stopIt) -> int {
        const i64 last = stopIt == LLONG_MAX ? chunkEnd[s] : std::min<i64>(chunkEnd[s], stopIt + 1);
        return tr.read(chunkFirst[s], last);
    };
    int slot = 0, pending = -1;
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
            const i64 stopHost = pinnedStop[pending];
            if ((rc = read_slot(pending, stopHost))) break;
            if (stopHost != LLONG_MAX) break;
        }
        if (enq >= imax) {
            if (pending != slot) {
                HS_HIP(hipEventSynchronize(evs[slot]));
                if ((rc = read_slot(slot, pinnedStop[slot]))) break;
            }
            break;
        }
        pending = slot;
        slot ^= 1;
    }
    if (rc) return rc;
    HS_HIP(hipStreamSynchronize(h->stream));
    RefState hst;
    HS_HIP(hipMemcpy(&hst, st, sizeof(RefState), hipMemcpyDeviceToHost));
    const i64 iters = (hst.stop == LLONG_MAX) ? enq : hst.stop + 1;
    if (iterations) *iterations = iters;
    if (delta_final) *delta_final = hst.delta[iters & 1];
    if (delta_zero) *delta_zero = hst.delta0;
    return HS_OK;
}

extern "C" int hs_cg_ref_order_f64(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N, int b,
                                   double eps, int64_t imax, int wg, int wgf, int fma, int64_t* iterations,
                                   double* delta_final, double* delta_zero) {
    return cg_ref_run(h, A, rhs, x, N, b, eps, imax, wg, wgf, fma, iterations, delta_final, delta_zero, nullptr);
}

extern "C" int hs_cg_ref_order_f64_trace(hs_handle_t h, const double* A, const double* rhs, double* x, int64_t N,
                                         int b, double eps, int64_t imax, int wg, int wgf, int fma,
                                         int64_t* iterations, double* delta_final, double* delta_zero,
                                         double* trace) {
    if (h && !trace) return -15;
    return cg_ref_run(h, A, rhs, x, N, b, eps, imax, wg, wgf, fma, iterations, delta_final, delta_zero, trace);
}
