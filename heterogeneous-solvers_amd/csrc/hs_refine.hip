// FP64 iterative refinement with a low-precision (FP32-stored) Cholesky
// factor, and CG preconditioned by that factor when refinement stalls. No
// counterpart in the reference.
//
// Every residual is r = y - K x on the FP64 K (read-once packed SYMV); the
// corrections come from the blocked triangular solves on the FP32 factor in
// FP64 arithmetic. Convergence is judged by the normwise backward error
//   omega = ||r||_inf / (||K||_inf ||x||_inf + ||y||_inf),
// with ||K||_inf from one pass of |K| times a vector of ones.
#include <algorithm>
#include <cmath>
#include <vector>

#include "hs_internal.h"

int hs_symv_into(hs_handle_t h, const double* A, const double* x, double* y, const double* rhs, i64 N, int b,
                 double* rowpart, double* colpart, double* part, int mode, bool abs_val);
int hs_launch_dot(hs_handle_t h, const double* x, const double* y, i64 begin, i64 len, double* part, int maxG,
                  double* out);

namespace {

constexpr int VT = 256;

__global__ void fill_ones_kernel(double* __restrict__ v, i64 rows, i64 N) {
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i < rows) v[i] = i < N ? 1.0 : 0.0;
}

__global__ void copy_kernel(const double* __restrict__ a, double* __restrict__ c, i64 rows) {
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i < rows) c[i] = a[i];
}

// y += alpha * x
__global__ void axpy_kernel(double* __restrict__ y, const double* __restrict__ x, double alpha, i64 rows) {
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i < rows) y[i] = y[i] + alpha * x[i];
}

// p = d + beta * p
__global__ void xpby_kernel(const double* __restrict__ d, double* __restrict__ p, double beta, i64 rows) {
    const i64 i = (i64)blockIdx.x * VT + threadIdx.x;
    if (i < rows) p[i] = d[i] + beta * p[i];
}

// per-work-group max |u| and max |v| (order-independent, so deterministic)
__global__ __launch_bounds__(VT) void maxabs2_kernel(const double* __restrict__ u, const double* __restrict__ v,
                                                      i64 rows, double* __restrict__ part) {
    __shared__ double ru[VT], rv[VT];
    double mu = 0.0, mv = 0.0;
    for (i64 i = (i64)blockIdx.x * VT + threadIdx.x; i < rows; i += (i64)gridDim.x * VT) {
        mu = fmax(mu, fabs(u[i]));
        mv = fmax(mv, fabs(v[i]));
    }
    ru[threadIdx.x] = mu;
    rv[threadIdx.x] = mv;
    __syncthreads();
    for (int s = VT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            ru[threadIdx.x] = fmax(ru[threadIdx.x], ru[threadIdx.x + s]);
            rv[threadIdx.x] = fmax(rv[threadIdx.x], rv[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ru[0];
        part[2 * blockIdx.x + 1] = rv[0];
    }
}

__global__ __launch_bounds__(VT) void maxabs2_final_kernel(const double* __restrict__ part, int G,
                                                            double* __restrict__ out) {
    __shared__ double ru[VT], rv[VT];
    double mu = 0.0, mv = 0.0;
    for (int g = threadIdx.x; g < G; g += VT) {
        mu = fmax(mu, part[2 * g]);
        mv = fmax(mv, part[2 * g + 1]);
    }
    ru[threadIdx.x] = mu;
    rv[threadIdx.x] = mv;
    __syncthreads();
    for (int s = VT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            ru[threadIdx.x] = fmax(ru[threadIdx.x], ru[threadIdx.x + s]);
            rv[threadIdx.x] = fmax(rv[threadIdx.x], rv[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = ru[0];
        out[1] = rv[0];
    }
}

struct Ctx {
    hs_handle_t h;
    const double* K;
    const float* L;
    i64 N;
    int b, n;
    i64 rows;
    double *Dinv, *rowpart, *colpart, *r, *d, *z, *p, *q, *part, *scal;
    double* pinned;
    dim3 vg, vb;

    int solve(double* v) { return hs_tri_solve_prepared<float>(h, L, Dinv, v, z, N, b); }
    int residual(const double* y, const double* x) {  // r = y - K x
        return hs_symv_into(h, K, x, r, y, N, b, rowpart, colpart, part, 2, false);
    }
    int maxabs2(const double* u, const double* v, double* mu, double* mv) {
        const int G = (int)std::min<i64>(1024, (rows + VT - 1) / VT);
        maxabs2_kernel<<<dim3((unsigned)G), vb, 0, h->stream>>>(u, v, rows, part);
        HS_LAUNCH_CHECK();
        maxabs2_final_kernel<<<1, vb, 0, h->stream>>>(part, G, scal);
        HS_LAUNCH_CHECK();
        HS_HIP(hipMemcpyAsync(pinned, scal, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HS_HIP(hipStreamSynchronize(h->stream));
        *mu = pinned[0];
        *mv = pinned[1];
        return HS_OK;
    }
    int dot(const double* u, const double* v, double* out) {
        int rc = hs_launch_dot(h, u, v, 0, rows, part, 1024, scal + 2);
        if (rc) return rc;
        HS_HIP(hipMemcpyAsync(pinned, scal + 2, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HS_HIP(hipStreamSynchronize(h->stream));
        *out = pinned[0];
        return HS_OK;
    }
    int copy(const double* a, double* c) {
        copy_kernel<<<vg, vb, 0, h->stream>>>(a, c, rows);
        HS_LAUNCH_CHECK();
        return HS_OK;
    }
    int axpy(double* y, const double* x, double a) {
        axpy_kernel<<<vg, vb, 0, h->stream>>>(y, x, a, rows);
        HS_LAUNCH_CHECK();
        return HS_OK;
    }
};

}  // namespace

size_t hs_ws_refine(i64 N, int b) {
    if (!hs_trsv_fast_ok(b)) return 0;
    const i64 n = (N + b - 1) / b, rows = n * b;
    const i64 nb = n * (n + 1) / 2;
    const i64 G = rows / 64 + 2048 + 64;
    return sizeof(double) * (size_t)(hs_dinv_doubles((int)n, b) + 2 * nb * b + 5 * rows + G + 16);
}

// The refinement loop behind both entries. Targets: the normwise backward
// error omega <= tol (skipped when tol == 0) and the relative residual
// ||y - K x||_2 / ||y||_2 <= rrtol (skipped when rrtol == 0); both must hold.
// Progress is measured as e = max(omega / tol, rr / rrtol) over the active
// targets: a refinement step that does not halve e hands over to PCG.
static int refine(hs_handle_t h, const double* K, const float* L32, const double* y, double* x, int64_t N, int b,
                  double tol, double rrtol, int max_ir, int max_pcg, hs_refine_report* rep) {
    Ctx c;
    c.h = h;
    c.K = K;
    c.L = L32;
    c.N = N;
    c.b = b;
    c.n = (int)((N + b - 1) / b);
    c.rows = (i64)c.n * b;
    const i64 nb = (i64)c.n * (c.n + 1) / 2;
    const i64 G = c.rows / 64 + 2048 + 64;
    const i64 nd = hs_dinv_doubles(c.n, b);
    double* w = nullptr;
    int rc = hs_work(h, hs_ws_refine(N, b), (void**)&w);
    if (rc) return rc;
    c.Dinv = w;
    w += nd;
    c.rowpart = w;
    w += nb * b;
    c.colpart = w;
    w += nb * b;
    c.r = w;
    w += c.rows;
    c.d = w;
    w += c.rows;
    c.z = w;
    w += c.rows;
    c.p = w;
    w += c.rows;
    c.q = w;
    w += c.rows;
    c.part = w;
    w += G;
    c.scal = w;
    c.pinned = (double*)h->pinned;
    c.vg = dim3((unsigned)((c.rows + VT - 1) / VT));
    c.vb = dim3(VT);
    *rep = hs_refine_report{0, 0, 0, 0, 0.0, 0.0, 0.0};

    if ((rc = hs_diag_inverse<float>(h, L32, N, b, c.Dinv))) return rc;
    // ||K||_inf = max row sum of |K|, ||y||_inf, ||y||_2
    fill_ones_kernel<<<c.vg, c.vb, 0, h->stream>>>(c.d, c.rows, N);
    HS_LAUNCH_CHECK();
    if ((rc = hs_symv_into(h, K, c.d, c.q, nullptr, N, b, c.rowpart, c.colpart, c.part, 0, true))) return rc;
    double knorm = 0.0, ynorm = 0.0, yy = 0.0;
    if ((rc = c.maxabs2(c.q, y, &knorm, &ynorm)) || (rc = c.dot(y, y, &yy))) return rc;
    const double ynorm2 = std::sqrt(yy);
    rep->knorm_inf = knorm;

    // the state of the current residual r = y - K x: omega, rr, and the
    // progress measure e (<= 1 when every active target holds)
    double omega = 0.0, rr = 0.0;
    auto measure = [&](double* e) -> int {
        double rmax, xmax, r2 = 0.0;
        int e1 = c.maxabs2(c.r, x, &rmax, &xmax);
        if (e1) return e1;
        omega = rmax / (knorm * xmax + ynorm);
        if (rrtol > 0.0 && (e1 = c.dot(c.r, c.r, &r2))) return e1;
        rr = ynorm2 > 0.0 ? std::sqrt(r2) / ynorm2 : 0.0;
        *e = std::max(tol > 0.0 ? omega / tol : 0.0, rrtol > 0.0 ? rr / rrtol : 0.0);
        return HS_OK;
    };

    // x0 = (L L^T)^{-1} y
    if ((rc = c.copy(y, x)) || (rc = c.solve(x))) return rc;
    double e = 0.0, ePrev = 0.0;
    bool converged = false, stalled = false;
    for (int it = 0;; ++it) {
        if ((rc = c.residual(y, x)) || (rc = measure(&e))) return rc;
        if (e <= 1.0) {
            converged = true;
            break;
        }
        if (it >= 1 && e > 0.5 * ePrev) {
            stalled = true;
            break;
        }
        if (it >= max_ir) break;
        ePrev = e;
        if ((rc = c.copy(c.r, c.d)) || (rc = c.solve(c.d)) || (rc = c.axpy(x, c.d, 1.0))) return rc;
        rep->ir_iterations = it + 1;
    }
    if (!converged && max_pcg > 0) {
        // r holds y - K x for the current x
        rep->used_pcg = stalled || rep->ir_iterations >= max_ir ? 1 : 0;
        double rho, pq;
        if ((rc = c.copy(c.r, c.d)) || (rc = c.solve(c.d)) || (rc = c.copy(c.d, c.p)) || (rc = c.dot(c.r, c.d, &rho)))
            return rc;
        for (int k = 1; k <= max_pcg; ++k) {
            if ((rc = hs_symv_into(h, K, c.p, c.q, nullptr, N, b, c.rowpart, c.colpart, c.part, 0, false))) return rc;
            if ((rc = c.dot(c.p, c.q, &pq))) return rc;
            const double alpha = rho / pq;
            if ((rc = c.axpy(x, c.p, alpha)) || (rc = c.axpy(c.r, c.q, -alpha))) return rc;
            rep->pcg_iterations = k;
            if ((rc = measure(&e))) return rc;
            if (e <= 1.0) {
                // confirm on the true residual before stopping
                if ((rc = c.residual(y, x)) || (rc = measure(&e))) return rc;
                if (e <= 1.0) {
                    converged = true;
                    break;
                }
            }
            double rhoNew;
            if ((rc = c.copy(c.r, c.d)) || (rc = c.solve(c.d)) || (rc = c.dot(c.r, c.d, &rhoNew))) return rc;
            xpby_kernel<<<c.vg, c.vb, 0, h->stream>>>(c.d, c.p, rhoNew / rho, c.rows);
            HS_LAUNCH_CHECK();
            rho = rhoNew;
        }
        if (!converged) {
            if ((rc = c.residual(y, x)) || (rc = measure(&e))) return rc;
        }
    }
    double r2 = 0.0;
    if ((rc = c.dot(c.r, c.r, &r2))) return rc;
    rep->backward_error = omega;
    rep->relative_residual = ynorm2 > 0.0 ? std::sqrt(r2) / ynorm2 : 0.0;
    rep->converged = e <= 1.0 ? 1 : 0;
    if (!rep->converged) {
        hs_set_error("hs_refine_solve: target not reached (backward error " + std::to_string(omega) +
                     ", relative residual " + std::to_string(rep->relative_residual) + ")");
        return HS_ERR_NOT_CONVERGED;
    }
    return HS_OK;
}

static int refine_args(hs_handle_t h, const double* K, const float* L32, const double* y, double* x, int64_t N,
                       int b, int max_ir, int max_pcg, hs_refine_report* rep) {
    if (!h) return -1;
    if (!K) return -2;
    if (!L32) return -3;
    if (!y) return -4;
    if (!x) return -5;
    if (N <= 0) return -6;
    if (!hs_trsv_fast_ok(b)) {
        hs_set_error("hs_refine_solve: matrix_bsz must be 64, 128 or 256");
        return -7;
    }
    if (max_ir < 0) return -9;
    if (max_pcg < 0) return -10;
    if (!rep) return -11;
    return HS_OK;
}

extern "C" int hs_refine_solve(hs_handle_t h, const double* K, const float* L32, const double* y, double* x,
                               int64_t N, int b, double tol, int max_ir, int max_pcg, hs_refine_report* rep) {
    int rc = refine_args(h, K, L32, y, x, N, b, max_ir, max_pcg, rep);
    if (rc) return rc;
    if (!(tol > 0.0)) return -8;
    return refine(h, K, L32, y, x, N, b, tol, 0.0, max_ir, max_pcg, rep);
}

extern "C" int hs_refine_solve_to(hs_handle_t h, const double* K, const float* L32, const double* y, double* x,
                                  int64_t N, int b, double tol, double rel_res_tol, int max_ir, int max_pcg,
                                  hs_refine_report* rep) {
    int rc = refine_args(h, K, L32, y, x, N, b, max_ir, max_pcg, rep);
    if (rc) return rc;
    if (!(tol >= 0.0) || !(rel_res_tol >= 0.0) || (tol == 0.0 && rel_res_tol == 0.0)) return -8;
    return refine(h, K, L32, y, x, N, b, tol, rel_res_tol, max_ir, max_pcg, rep);
}
