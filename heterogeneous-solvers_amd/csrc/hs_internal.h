// Internal declarations shared by the libhs_amd.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "hs_abi.h"

typedef int64_t i64;

// The handle's device workspace is sized by hs_workspace_reserve (a set-up
// call); solver entries take their scratch from it through hs_work and never
// allocate.
constexpr int HS_HANDLE_EVENTS = 8;  // ordering events created with the handle

struct hs_handle_s {
    int device = 0;
    hipStream_t stream = nullptr;   // the handle's ordering stream
    hipStream_t stream2 = nullptr;  // look-ahead panel stream (same priority)
    std::vector<hipEvent_t> events; // reusable ordering events (no timing), HS_HANDLE_EVENTS at creation
    int* info_dev = nullptr;   // sticky first bad pivot (1-based global row), 0 = none
    // device workspace for partial sums, vectors and scalars of the entries
    void* work = nullptr;
    size_t work_bytes = 0;
    // pinned host word used to poll device-side loop state without a full copy
    void* pinned = nullptr;
    int cu_count = 256;
    // inverses of the 16x16 diagonal sub-blocks of the last factored diagonal
    // block (up to 16 of them, b <= 256), written by the panel POTRF, read by
    // the panel TRSM that follows it on the same stream
    double* trsm_w = nullptr;
    // the fp16 hi+lo split's scale of the mixed factor in progress: s, 1/s^2
    // and the bound they come from (HS_MIXED_F16X3)
    float* hilo_scale = nullptr;
};

struct hs_event_s {
    hipEvent_t ev = nullptr;
};

void hs_set_error(const std::string& msg);
int hs_hip_fail(hipError_t e, const char* what);
// The first `bytes` of the handle workspace in *ptr, or HS_ERR_WORKSPACE when
// hs_workspace_reserve has reserved less. Never allocates.
int hs_work(hs_handle_t h, size_t bytes, void** ptr);
// Workspace bytes each entry family needs for a system of order N in blocks
// of b (M test points for the GP prediction); hs_workspace_query takes the
// maximum of them.
size_t hs_ws_symv(i64 N, int b);
size_t hs_ws_cg(i64 N, int b);
size_t hs_ws_cg_ref(i64 N, int b, int wg);
size_t hs_ws_dot_ref(i64 len, int wg);
size_t hs_ws_dot();
size_t hs_ws_tri_solve(i64 N, int b);
size_t hs_ws_refine(i64 N, int b);
size_t hs_ws_mixed(i64 N, int b);
size_t hs_ws_adjoint(i64 N, int b);
size_t hs_ws_gp_predict(i64 N, i64 M);
size_t hs_ws_rbf_window(i64 N);
// Looks up the allocation-free CG's scratch address for `device` (current).
int hs_noalloc_setup(int device);
// Ensures the handle's event pool holds at least `count` events.
int hs_ensure_events(hs_handle_t h, size_t count);

#define HS_HIP(call)                                                \
    do {                                                            \
        hipError_t _e = (call);                                     \
        if (_e != hipSuccess) return hs_hip_fail(_e, #call);        \
    } while (0)

#define HS_LAUNCH_CHECK() HS_HIP(hipGetLastError())

// Per-iteration timing of the device-resident CG loops (the traced entries):
// timing events at five marks of every iteration (before the mat-vec, after
// it, after alpha and the x update, after the residual, after delta and the d
// update), held for the two chunks of iterations in flight and read back when
// the host sees a chunk complete. Iteration i yields out[5 i .. 5 i + 4]:
// the four phase times and the whole iteration, in ms.
struct IterTrace {
    static constexpr int M = 5;
    double* out = nullptr;
    i64 cap = 0;
    int chunk = 1;
    std::vector<hipEvent_t> ev;
    int start(double* o, i64 capacity, int chunkIters) {
        out = o;
        cap = capacity;
        chunk = chunkIters;
        if (!out) return HS_OK;
        ev.assign((size_t)2 * chunk * M, nullptr);
        for (auto& e : ev) HS_HIP(hipEventCreate(&e));
        return HS_OK;
    }
    ~IterTrace() {
        for (auto e : ev)
            if (e) hipEventDestroy(e);
    }
    hipEvent_t at(i64 it, int m) const { return ev[(size_t)((((it / chunk) & 1) * chunk + it % chunk) * M + m)]; }
    int mark(hipStream_t s, i64 it, int m) {
        if (!out) return HS_OK;
        HS_HIP(hipEventRecord(at(it, m), s));
        return HS_OK;
    }
    // iterations [first, last) have completed on the device
    int read(i64 first, i64 last) {
        if (!out) return HS_OK;
        for (i64 it = first; it < last && it < cap; ++it) {
            float t = 0.f;
            for (int m = 0; m + 1 < M; ++m) {
                HS_HIP(hipEventElapsedTime(&t, at(it, m), at(it, m + 1)));
                out[it * M + m] = t;
            }
            HS_HIP(hipEventElapsedTime(&t, at(it, 0), at(it, M - 1)));
            out[it * M + M - 1] = t;
        }
        return HS_OK;
    }
};

// packed block ID of (I,J), I >= J
__host__ __device__ inline i64 hs_block_id(i64 I, i64 J, i64 n) { return J * n - J * (J - 1) / 2 + (I - J); }

// Row-by-row decode of a linear index t over a lower triangle with diagonal:
// t = r(r+1)/2 + c, 0 <= c <= r.
__host__ __device__ inline void hs_tri_decode(i64 t, int* r, int* c) {
    i64 rr = (i64)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while (rr * (rr + 1) / 2 > t) --rr;
    while ((rr + 1) * (rr + 2) / 2 <= t) ++rr;
    *r = (int)rr;
    *c = (int)(t - rr * (rr + 1) / 2);
}

// kernels implemented in other translation units, used by the solver drivers
int hs_launch_potrf(hs_handle_t h, double* A, i64 N, int b, i64 blockID, int blockRow);
int hs_launch_trsm(hs_handle_t h, double* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount);
// trailing update over rows [rowStart, rowStart+rowCount) of step k;
// mode 1 = diagonal blocks only (SYRK), 2 = off-diagonal only (GEMM), 3 = both
int hs_launch_trailing(hs_handle_t h, double* A, int b, int k, int rowStart, int rowCount, int n, int mode);
// same on an explicit stream with np panels k .. k+np-1; modes 4 (the
// triangle from row k + 2 np), 5 (block column k+1) and 6 (block columns
// k+np, k+np+1) split a step for the look-ahead driver
// (halfCu: one 128-tile work-group per CU, leaving the other half of every CU
// to the panel chain's kernels)
int hs_launch_trailing_on(hs_handle_t h, hipStream_t stream, double* A, int b, int k, int rowStart, int rowCount, int n,
                          int mode, int np, bool halfCu = false);
// panel kernels for b % 32 == 0 (POTRF) / b % 16 == 0 (TRSM), b <= 256;
// W = the 16x16 diagonal sub-block inverses of the factor (POTRF writes them
// when W is non-null; a TRSM given W == nullptr computes them into trsm_w)
bool hs_potrf_fast_ok(int b);
bool hs_trsm_fast_ok(int b);
// (minLds: the launch asks for at least this much LDS, so that it only starts
// on a CU without other resident work)
template <typename ST>
int hs_launch_potrf_fast(hs_handle_t h, ST* A, i64 N, int b, i64 blockID, int blockRow, double* W, size_t minLds = 0);
template <typename ST>
int hs_launch_trsm_fast(hs_handle_t h, ST* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount,
                        const double* W);
// factorisation steps: the panel kernels where b allows them (the POTRF then
// leaves the sub-block inverses for the TRSM of the same column), else the
// element-wise kernels (FP64 storage only)
template <typename ST>
int hs_factor_potrf(hs_handle_t h, ST* A, i64 N, int b, i64 blockID, int blockRow, size_t minLds = 0);
template <typename ST>
int hs_factor_trsm(hs_handle_t h, ST* A, int b, i64 blockID, int blockRow, int blockStart, int blockCount);
// fast blocked triangular solves (b in {64, 128, 256}); factor in FP64 or FP32
bool hs_trsv_fast_ok(int b);
// doubles of the Dinv workspace of hs_diag_inverse for n blocks of b
i64 hs_dinv_doubles(int n, int b);
template <typename T>
int hs_diag_inverse(hs_handle_t h, const T* L, i64 N, int b, double* Dinv);
// phases: bit 0 forward (z = L^{-1} v), bit 1 backward (v = L^{-T} z)
template <typename T>
int hs_tri_solve_prepared(hs_handle_t h, const T* L, const double* Dinv, double* v, double* z, i64 N, int b,
                          int phases = 3);
// the reference-order mat-vec (1 <= b <= 256), y = A x over n block rows;
// fma selects contracted multiply-adds; stop / it gate it like the CG kernels
bool hs_symv_ordered_ok(int b);
int hs_launch_symv_ordered(hs_handle_t h, const double* A, const double* x, double* y, int n, int b, int fma,
                           const i64* stop, i64 it);
int hs_launch_trsv_block(hs_handle_t h, const double* A, double* x, i64 N, int b, int blockRow, i64 blockID, int transposed);
int hs_launch_col_update(hs_handle_t h, const double* A, double* x, int b, int blockStart, int blockCount, int blockRow,
                         i64 blockID, int n, int transposed);
