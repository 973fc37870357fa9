// Heterogeneous CPU + GPU solves at a GPU share below 1.0: the reference's row
// splits (CG.cpp:38-163, 277-537; Cholesky.cpp:56-259, 361-481) with the GPU
// half on the C ABI's per-op kernels and the CPU half on this library's own
// OpenMP kernels (cpu_kernels.cpp), exchanging vector parts and matrix blocks
// with synchronous copies. The split is the load balancer's: the static one
// keeps it; a balancer that changes the share moves the boundary.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>

#include "cpu_kernels.hpp"
#include "hs/hs.hpp"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

// A pair of HIP timing events on the handle's stream, created at first use
// (a CPU-only solve never touches the device).
struct GpuSpan {
    hs_event_t a = nullptr, b = nullptr;
    bool used = false;
    ~GpuSpan() {
        if (a) hs_event_destroy(a);
        if (b) hs_event_destroy(b);
    }
    void start(hs_handle_t h) {
        if (!a) {
            hs::check(hs_event_create(&a), "hs_event_create");
            hs::check(hs_event_create(&b), "hs_event_create");
        }
        hs::check(hs_event_record(h, a), "hs_event_record");
        used = true;
    }
    void stop(hs_handle_t h) { hs::check(hs_event_record(h, b), "hs_event_record"); }
    double ms() {
        if (!used) return 0.0;
        float v = 0.f;
        hs::check(hs_event_elapsed_ms(a, b, &v), "hs_event_elapsed_ms");
        used = false;
        return v;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// CG: block rows [0, nG) on the GPU, [nG, n) on the CPU, nG = ceil(n p)
// (CG.cpp:43-46). Both sides hold the whole matrix (the reference copies all of
// it to the GPU when it fits, CG.cpp:165-234); each mat-vec needs the whole d,
// so the two parts of d are exchanged before it (CG.cpp:345-353), and of x
// before a true residual (:446-455). Dot products are the two sides' partial
// sums added on the host.
// ---------------------------------------------------------------------------
void CG::solveSplit() {
    const auto start = clk::now();
    metricsTracker.startTracking();
    const int n = A.blockCountXY, bs = A.blockSize;
    const std::size_t len = static_cast<std::size_t>(n) * bs;
    const int64_t N = static_cast<int64_t>(A.N);
    double share = loadBalancer->currentProportionGPU;
    int nG = static_cast<int>(std::ceil(static_cast<double>(n) * share));
    const bool useGPU = share > 0.0;
    hs_handle_t h = useGPU ? gpuQueue.handle() : nullptr;

    auto t0 = clk::now();
    hs::DeviceArray<double> dA, db, dx, dr, dd, dq, ds;
    if (useGPU) {
        gpuQueue.reserve(A.N, bs);
        dA.resize(A.matrixData.size());
        db.resize(len);
        dx.resize(len);
        dr.resize(len);
        dd.resize(len);
        dq.resize(len);
        ds.resize(1);
        hs::check(hs_memcpy_htod(h, dA.data(), A.matrixData.data(), A.matrixData.size() * 8), "hs_memcpy_htod");
        hs::check(hs_memcpy_htod(h, db.data(), b.rightHandSideData.data(), len * 8), "hs_memcpy_htod");
    }
    std::vector<double> r(len, 0.0), d(len, 0.0), q(len, 0.0);
    x.assign(len, 0.0);
    metricsTracker.memoryInitTime = ms_since(t0);

    const double* Ah = A.matrixData.data();
    const double* bh = b.rightHandSideData.data();
    auto gRows = [&]() { return static_cast<std::size_t>(nG) * bs; };
    auto gpuDot = [&](const double* u, const double* v) -> double {
        if (nG == 0) return 0.0;
        double s = 0.0;
        hs::check(hs_dot_f64(h, u, v, bs, 0, nG, ds.data()), "hs_dot_f64");
        hs::check(hs_memcpy_dtoh(h, &s, ds.data(), 8), "hs_memcpy_dtoh");
        return s;
    };
    // the host and device copies of a vector agree on both parts afterwards
    auto exchange = [&](double* host, double* dev) {
        if (!useGPU) return;
        const std::size_t g = gRows();
        if (g < len) hs::check(hs_memcpy_htod(h, dev + g, host + g, (len - g) * 8), "hs_memcpy_htod");
        if (g > 0) hs::check(hs_memcpy_dtoh(h, host, dev, g * 8), "hs_memcpy_dtoh");
    };
    // r = b - A x on both parts (x already exchanged)
    auto residual = [&]() {
        if (nG > 0) {
            hs::check(hs_symv_rows_f64(h, dA.data(), dx.data(), dr.data(), N, bs, nG), "hs_symv_rows_f64");
            hs::check(hs_vec_sub_f64(h, db.data(), dr.data(), dr.data(), bs, 0, nG), "hs_vec_sub_f64");
        }
        if (nG < n) {
            hs::cpu::symv_rows(Ah, x.data(), r.data(), bs, nG, n - nG, n);
            hs::cpu::sub(bh, r.data(), r.data(), gRows(), len);
        }
        if (useGPU) hs::check(hs_sync(h), "hs_sync");
    };

    // initCG (CG.cpp:277-340): x = 0, r = b - A x, d = r, delta_0 = r.r
    if (useGPU) hs::check(hs_memset_zero(h, dx.data(), len * 8), "hs_memset_zero");
    residual();
    if (nG > 0) hs::check(hs_memcpy_dtod(h, dd.data(), dr.data(), gRows() * 8), "hs_memcpy_dtod");
    std::copy(r.begin() + gRows(), r.end(), d.begin() + gRows());
    double deltaNew = gpuDot(dr.data(), dr.data()) + hs::cpu::dot(r.data(), r.data(), gRows(), len);
    const double delta0 = deltaNew;
    const double eps2 = conf::epsilon * conf::epsilon;

    GpuSpan symvSpan;
    std::size_t it = 0;
    while (it < conf::iMax && deltaNew > eps2 * delta0) {
        const auto tIt = clk::now();
        if (it % loadBalancer->updateInterval == 0 && it != 0) {
            // a new share moves the boundary: both copies of x, r and d are
            // made whole first, so either side may take over rows
            const double s2 = loadBalancer->getNewProportionGPU(metricsTracker);
            const int nG2 = useGPU ? std::min(n, static_cast<int>(std::ceil(static_cast<double>(n) * s2))) : 0;
            if (nG2 != nG) {
                exchange(x.data(), dx.data());
                exchange(r.data(), dr.data());
                exchange(d.data(), dd.data());
                nG = nG2;
                share = s2;
            }
        }
        // q = A d (CG.cpp:342-398)
        const auto tq = clk::now();
        exchange(d.data(), dd.data());
        const double memcopyMs = ms_since(tq);
        double cpuSymvMs = 0.0;
        if (nG > 0) {
            symvSpan.start(h);
            hs::check(hs_symv_rows_f64(h, dA.data(), dd.data(), dq.data(), N, bs, nG), "hs_symv_rows_f64");
            symvSpan.stop(h);
        }
        if (nG < n) {
            const auto tc = clk::now();
            hs::cpu::symv_rows(Ah, d.data(), q.data(), bs, nG, n - nG, n);
            cpuSymvMs = ms_since(tc);
        }
        if (useGPU) hs::check(hs_sync(h), "hs_sync");
        const double gpuSymvMs = symvSpan.ms();
        const auto t2 = clk::now();
        // alpha = delta_new / d.q
        const double alpha = deltaNew / (gpuDot(dd.data(), dq.data()) + hs::cpu::dot(d.data(), q.data(), gRows(), len));
        const auto t3 = clk::now();
        // x = x + alpha d
        if (nG > 0) hs::check(hs_vec_scale_add_f64(h, dx.data(), alpha, dd.data(), dx.data(), bs, 0, nG), "x");
        hs::cpu::scale_add(x.data(), alpha, d.data(), x.data(), gRows(), len);
        if (useGPU) hs::check(hs_sync(h), "hs_sync");
        const auto t4 = clk::now();
        if (it % 50 == 0) {
            exchange(x.data(), dx.data());
            residual();
        } else {
            if (nG > 0) hs::check(hs_vec_scale_add_f64(h, dr.data(), -1.0 * alpha, dq.data(), dr.data(), bs, 0, nG), "r");
            hs::cpu::scale_add(r.data(), -1.0 * alpha, q.data(), r.data(), gRows(), len);
        }
        const auto t5 = clk::now();
        const double deltaOld = deltaNew;
        deltaNew = gpuDot(dr.data(), dr.data()) + hs::cpu::dot(r.data(), r.data(), gRows(), len);
        const auto t6 = clk::now();
        const double beta = deltaNew / deltaOld;
        if (nG > 0) hs::check(hs_vec_scale_add_f64(h, dr.data(), beta, dd.data(), dd.data(), bs, 0, nG), "d");
        hs::cpu::scale_add(r.data(), beta, d.data(), d.data(), gRows(), len);
        if (useGPU) hs::check(hs_sync(h), "hs_sync");
        const auto tEnd = clk::now();
        auto span = [](clk::time_point a, clk::time_point c) {
            return std::chrono::duration<double, std::milli>(c - a).count();
        };
        metricsTracker.updateMetrics(it, nG, n - nG, span(tIt, tEnd), loadBalancer->updateInterval);
        metricsTracker.memcopy_d.push_back(memcopyMs);
        metricsTracker.matrixVectorTimes_GPU.push_back(gpuSymvMs);
        metricsTracker.matrixVectorTimes_CPU.push_back(cpuSymvMs);
        metricsTracker.times_q.push_back(span(tq, t2));
        metricsTracker.times_alpha.push_back(span(t2, t3));
        metricsTracker.times_x.push_back(span(t3, t4));
        metricsTracker.times_r.push_back(span(t4, t5));
        metricsTracker.times_delta.push_back(span(t5, t6));
        metricsTracker.times_d.push_back(span(t6, tEnd));
        ++it;
    }
    iterations = it;
    deltaZero = delta0;
    deltaFinal = deltaNew;
    t0 = clk::now();
    if (nG > 0) hs::check(hs_memcpy_dtoh(h, x.data(), dx.data(), gRows() * 8), "hs_memcpy_dtoh");
    metricsTracker.resultCopyTime = ms_since(t0);
    metricsTracker.totalTime = ms_since(start);
    loadBalancer->currentProportionGPU = share;
}

// ---------------------------------------------------------------------------
// Cholesky: per column k the trailing block rows split into a CPU part
// [k+1, sG) and a GPU part [sG, n) with nG = min(max(ceil(count p), minG),
// count) GPU rows of the count = n-k-1 below the diagonal
// (Cholesky.cpp:56-157); the GPU keeps at least conf::minBlockCountCholesky
// rows. Both sides hold the whole matrix; a row belongs to the side of its
// part, and when the boundary moves the rows that change sides are copied
// across. The diagonal block is factored where its row lives and copied to the
// other side (:159-184); the CPU's panel blocks are copied to the GPU after the
// TRSM (:186-207), which the GPU's updates of its rows read.
// ---------------------------------------------------------------------------
void Cholesky::solveSplit() {
    const auto start = clk::now();
    metricsTracker.startTracking();
    const int n = A.blockCountXY, bs = A.blockSize;
    const std::size_t bb = static_cast<std::size_t>(bs) * bs;
    const int64_t N = static_cast<int64_t>(A.N);
    double share = conf::initialProportionGPU;
    const bool useGPU = share > 0.0;
    const int minG = useGPU ? conf::minBlockCountCholesky : 0;
    hs_handle_t h = useGPU ? gpuQueue.handle() : nullptr;
    double* Ah = A.matrixData.data();
    const std::size_t bytes = A.matrixData.size() * sizeof(double);

    auto t0 = clk::now();
    if (useGPU) {
        gpuQueue.reserve(A.N, bs);
        if (!A_gpu) {
            void* p = nullptr;
            hs::check(hs_malloc(&p, bytes), "hs_malloc");
            A_gpu = static_cast<double*>(p);
        }
        hs::check(hs_memcpy_htod(h, A_gpu, Ah, bytes), "hs_memcpy_htod");
        hs::check(hs_reset_info(h), "hs_reset_info");
    }
    metricsTracker.memoryInitTime = ms_since(t0);

    // rows [r0, r1) of every column J < r1, host <-> device (each column's
    // blocks of consecutive rows are contiguous)
    auto copyRows = [&](int r0, int r1, bool toHost) {
        for (int J = 0; J < r1; ++J) {
            const int lo = std::max(r0, J);
            if (lo >= r1) continue;
            const std::size_t off = hs::cpu::blockID(lo, J, n) * bb, cnt = static_cast<std::size_t>(r1 - lo) * bb;
            if (toHost) hs::check(hs_memcpy_dtoh(h, Ah + off, A_gpu + off, cnt * 8), "hs_memcpy_dtoh");
            else hs::check(hs_memcpy_htod(h, A_gpu + off, Ah + off, cnt * 8), "hs_memcpy_htod");
        }
    };

    int nG = static_cast<int>(std::ceil(static_cast<double>(n - 1) * share));
    int nC = n - 1 - nG;
    int sG = nC + 1;
    int offsetGemm = 0;
    int64_t cpuInfo = 0;
    GpuSpan sTrsm, sSyrk, sGemm;
    for (int k = 0; k < n; ++k) {
        const auto tCol = clk::now();
        const std::size_t id = hs::cpu::blockID(k, k, n);
        // the split of this column (shiftSplit, Cholesky.cpp:56-157)
        const auto tShift = clk::now();
        if (k % loadBalancer->updateInterval == 0 && k != 0 && share != 0.0 && share != 1.0)
            share = loadBalancer->getNewProportionGPU(metricsTracker);
        const int count = n - (k + 1);
        const int nGnew = std::min(std::max(static_cast<int>(std::ceil(count * share)), minG), count);
        const int nCnew = std::max(count - nGnew, 0);
        const int sGnew = std::max(n - nGnew, k + 1);
        if (useGPU && sGnew > sG) copyRows(sG, sGnew, true);   // GPU rows move to the CPU
        if (useGPU && sGnew < sG) copyRows(sGnew, sG, false);  // CPU rows move to the GPU
        nG = nGnew;
        nC = nCnew;
        sG = sGnew;
        const double shiftMs = ms_since(tShift);
        if (sG <= k + 1) offsetGemm = 1;

        // POTRF where row k lives (Cholesky.cpp:159-184)
        const auto tPotrf = clk::now();
        const bool cpuHolds = (k < n - std::max(nG, minG) && share != 1.0) || !useGPU;
        if (cpuHolds) {
            const int64_t info = hs::cpu::potrf(Ah, A.N, bs, id, k);
            if (info && !cpuInfo) cpuInfo = info;
            if (useGPU) hs::check(hs_memcpy_htod(h, A_gpu + id * bb, Ah + id * bb, bb * 8), "hs_memcpy_htod");
        } else {
            hs::check(hs_potrf_diag_f64(h, A_gpu, N, bs, static_cast<int>(id), k), "hs_potrf_diag_f64");
            hs::check(hs_memcpy_dtoh(h, Ah + id * bb, A_gpu + id * bb, bb * 8), "hs_memcpy_dtoh");
        }
        const double potrfMs = ms_since(tPotrf);

        // TRSM of the column, CPU rows [k+1, k+1+nC), GPU rows [sG, n)
        const auto tTrsm = clk::now();
        double cpuTrsm = 0.0, copyMs = 0.0;
        if (nG > 0) {
            sTrsm.start(h);
            hs::check(hs_trsm_panel_f64(h, A_gpu, bs, static_cast<int>(id), k, sG, nG), "hs_trsm_panel_f64");
            sTrsm.stop(h);
        }
        if (nC > 0) {
            const auto tc = clk::now();
            hs::cpu::trsm(Ah, bs, id, k, k + 1, nC);
            cpuTrsm = ms_since(tc);
        }
        if (useGPU) hs::check(hs_sync(h), "hs_sync");
        if (useGPU && nC > 0 && share != 1.0) {
            const auto tc = clk::now();
            hs::check(hs_memcpy_htod(h, A_gpu + (id + 1) * bb, Ah + (id + 1) * bb, static_cast<std::size_t>(nC) * bb * 8),
                      "hs_memcpy_htod");
            copyMs = ms_since(tc);
        }
        const double trsmMs = ms_since(tTrsm);

        // SYRK on the diagonal blocks below (Cholesky.cpp:209-221)
        const auto tSyrk = clk::now();
        double cpuSyrk = 0.0;
        if (nG > 0) {
            sSyrk.start(h);
            hs::check(hs_syrk_diag_f64(h, A_gpu, bs, static_cast<int>(id), k, sG, nG, n), "hs_syrk_diag_f64");
            sSyrk.stop(h);
        }
        if (nC > 0) {
            const auto tc = clk::now();
            hs::cpu::syrk(Ah, bs, k, k + 1, nC, n);
            cpuSyrk = ms_since(tc);
        }
        if (useGPU) hs::check(hs_sync(h), "hs_sync");
        const double syrkMs = ms_since(tSyrk);

        // GEMM on the off-diagonal blocks (Cholesky.cpp:223-259)
        const auto tGemm = clk::now();
        double cpuGemm = 0.0;
        if (nG > 1) {
            sGemm.start(h);
            hs::check(hs_gemm_trailing_f64(h, A_gpu, bs, static_cast<int>(id), k, sG + offsetGemm, nG - offsetGemm, n),
                      "hs_gemm_trailing_f64");
            sGemm.stop(h);
        }
        if (nC > 1) {
            const auto tc = clk::now();
            hs::cpu::gemm(Ah, bs, k, k + 2, nC - 1, n);
            cpuGemm = ms_since(tc);
        }
        if (useGPU) hs::check(hs_sync(h), "hs_sync");
        const double gemmMs = ms_since(tGemm);

        // the reference's per-column record (Cholesky.cpp:261-359)
        metricsTracker.shiftTimes.push_back(shiftMs);
        metricsTracker.choleskyDiagonalBlockTimes.push_back(potrfMs);
        metricsTracker.copyTimes.push_back(copyMs);
        metricsTracker.triangularSolveTimes_total.push_back(trsmMs);
        metricsTracker.triangularSolveTimes_CPU.push_back(cpuTrsm);
        metricsTracker.triangularSolveTimes_GPU.push_back(sTrsm.ms());
        metricsTracker.matrixMatrixDiagonalTimes_total.push_back(syrkMs);
        metricsTracker.matrixMatrixDiagonalTimes_CPU.push_back(cpuSyrk);
        metricsTracker.matrixMatrixDiagonalTimes_GPU.push_back(sSyrk.ms());
        metricsTracker.matrixMatrixTimes_total.push_back(gemmMs);
        metricsTracker.matrixMatrixTimes_CPU.push_back(cpuGemm);
        metricsTracker.matrixMatrixTimes_GPU.push_back(sGemm.ms());
        metricsTracker.updateMetrics(k, nG, nC, ms_since(tCol), loadBalancer->updateInterval);
    }
    // rows still on the GPU at the end come back (copyResultFromGPU, :361-397)
    t0 = clk::now();
    int64_t gpuInfo = 0;
    if (useGPU) {
        copyRows(std::min(sG, n - std::min(minG, n)), n, true);
        hs::check(hs_get_info(h, &gpuInfo), "hs_get_info");
    }
    metricsTracker.resultCopyTime = ms_since(t0);
    info = cpuInfo && gpuInfo ? std::min(cpuInfo, gpuInfo) : (cpuInfo ? cpuInfo : gpuInfo);
    if (info > 0) std::cerr << "warning: non-positive pivot at row " << info << " (matrix not SPD)" << std::endl;
    metricsTracker.totalTime = ms_since(start);
    loadBalancer->currentProportionGPU = share;
}
