// CG, Cholesky and TriangularSystemSolver of the host API over the C ABI.
// At GPU share 1.0 the whole solve runs on the device; at a share below 1.0
// the rows are split between the CPU and the GPU as the reference splits them
// (hetero.cpp), and the triangular solves run on the CPU
// (TriangularSystemSolver.cpp:18).
#include "cpu_kernels.hpp"
#include <chrono>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "hs/hs.hpp"

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0) { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }

std::string timeString() {
    std::time_t t = std::time(nullptr);
    std::tm tm = *std::localtime(&t);
    std::ostringstream s;
    s << std::put_time(&tm, "%Y_%m_%d__%H_%M_%S");
    return s.str();
}

void requireValidShare(double share, const char* who) {
    if (!(share >= 0.0 && share <= 1.0))
        throw std::invalid_argument(std::string(who) + ": GPU share must lie in [0, 1], got " + std::to_string(share));
}

// UtilityFunctions::writeResult (UtilityFunctions.cpp:19-33): one line, values
// separated by spaces, fixed notation with 20 digits
template <typename V>
void writeResult(const std::string& dir, const V& x) {
    std::ofstream out(dir + "/x_result.txt");
    out << std::setprecision(20) << std::fixed;
    for (std::size_t i = 0; i < x.size(); ++i) out << x[i] << " ";
    out << std::endl;
}

}  // namespace

CG::CG(SymmetricMatrix& A, RightHandSide& b, hs::Queue& cpuQueue, hs::Queue& gpuQueue,
       std::shared_ptr<LoadBalancer> loadBalancer)
    : A(A), b(b), cpuQueue(cpuQueue), gpuQueue(gpuQueue), loadBalancer(std::move(loadBalancer)) {
    if (A.N != b.N) {
        throw std::invalid_argument("Dimensions of A and b do not match: " + std::to_string(A.N) +
                                    " != " + std::to_string(b.N));
    }
    if (A.blockCountXY != b.blockCountX) {
        throw std::invalid_argument("Block count of A and b do not match: " + std::to_string(A.blockCountXY) +
                                    " != " + std::to_string(b.blockCountX));
    }
    x.assign(b.rightHandSideData.size(), 0.0);
}

void CG::solveHeterogeneous() {
    requireValidShare(loadBalancer->currentProportionGPU, "CG");
    if (loadBalancer->currentProportionGPU < 1.0) {
        solveSplit();
        finish();
        return;
    }
    const auto start = clk::now();
    metricsTracker.startTracking();
    hs_handle_t h = gpuQueue.handle();
    const std::size_t bytesA = A.matrixData.size() * sizeof(double);
    const std::size_t bytesV = b.rightHandSideData.size() * sizeof(double);
    if (bytesA + 6 * bytesV > static_cast<std::size_t>(0.9 * static_cast<double>(gpuQueue.global_mem_size()))) {
        throw std::runtime_error("GPU memory not sufficient for the matrix (this build keeps it whole on the GPU)");
    }
    auto t0 = clk::now();
    gpuQueue.reserve(A.N, A.blockSize);
    hs::DeviceArray<double> dA(A.matrixData.size()), db(b.rightHandSideData.size()), dx(b.rightHandSideData.size());
    hs::check(hs_memcpy_htod(h, dA.data(), A.matrixData.data(), bytesA), "hs_memcpy_htod");
    hs::check(hs_memcpy_htod(h, db.data(), b.rightHandSideData.data(), bytesV), "hs_memcpy_htod");
    metricsTracker.memoryInitTime = ms_since(t0);

    t0 = clk::now();
    int64_t it = 0;
    double df = 0.0, d0 = 0.0;
    // per-iteration phase times from device events (5 per iteration)
    std::vector<double> trace(5 * static_cast<std::size_t>(conf::iMax), 0.0);
    if (conf::referenceOrder) {
        hs::check(hs_cg_ref_order_f64_trace(h, dA.data(), db.data(), dx.data(), static_cast<int64_t>(A.N),
                                            A.blockSize, conf::epsilon, static_cast<int64_t>(conf::iMax),
                                            conf::workGroupSizeVector, conf::workGroupSizeFinalScalarProduct,
                                            conf::referenceOrderFMA ? 1 : 0, &it, &df, &d0, trace.data()),
                  "hs_cg_ref_order_f64_trace");
    } else {
        hs::check(hs_cg_f64_trace(h, dA.data(), db.data(), dx.data(), static_cast<int64_t>(A.N), A.blockSize,
                                  conf::epsilon, static_cast<int64_t>(conf::iMax), &it, &df, &d0, trace.data()),
                  "hs_cg_f64_trace");
    }
    iterations = static_cast<std::size_t>(it);
    deltaZero = d0;
    deltaFinal = df;
    // The reference records per iteration (CG.cpp:109-129): the iteration
    // time with its block counts, the six step times and the mat-vec time.
    // Here alpha and delta are fused into the x and d kernels: their times
    // are inside times_x / times_d, and times_alpha / times_delta hold 0. At
    // share 1.0 no d exchange happens: memcopy_d holds 0.
    for (std::size_t i = 0; i < iterations; ++i) {
        const double* t = &trace[5 * i];
        metricsTracker.updateMetrics(i, A.blockCountXY, 0, t[4], loadBalancer->updateInterval);
        metricsTracker.times_q.push_back(t[0]);
        metricsTracker.times_alpha.push_back(0.0);
        metricsTracker.times_x.push_back(t[1]);
        metricsTracker.times_r.push_back(t[2]);
        metricsTracker.times_delta.push_back(0.0);
        metricsTracker.times_d.push_back(t[3]);
        metricsTracker.memcopy_d.push_back(0.0);
        metricsTracker.matrixVectorTimes_GPU.push_back(t[0]);
        metricsTracker.matrixVectorTimes_CPU.push_back(0.0);
    }

    t0 = clk::now();
    hs::check(hs_memcpy_dtoh(h, x.data(), dx.data(), bytesV), "hs_memcpy_dtoh");
    metricsTracker.resultCopyTime = ms_since(t0);
    metricsTracker.totalTime = ms_since(start);
    finish();
}

// CG.cpp:131-162: report, metrics.json under conf::outputPath, x_result.txt
void CG::finish() {
    std::cout << "Total time: " << metricsTracker.totalTime << "ms (" << iterations << " iterations)" << std::endl;
    metricsTracker.endTracking();
    std::string dir = conf::outputPath + "/" + timeString();
    std::filesystem::create_directories(dir);
    metricsTracker.writeJSON(dir);
    if (conf::writeResult) writeResult(".", x);
}

Cholesky::Cholesky(SymmetricMatrix& A, hs::Queue& cpuQueue, hs::Queue& gpuQueue,
                   std::shared_ptr<LoadBalancer> loadBalancer)
    : A(A), cpuQueue(cpuQueue), gpuQueue(gpuQueue), loadBalancer(std::move(loadBalancer)) {}

Cholesky::~Cholesky() {
    if (A_gpu) hs_free(A_gpu);
}

void Cholesky::solve_heterogeneous() {
    requireValidShare(conf::initialProportionGPU, "Cholesky");
    if (conf::initialProportionGPU < 1.0) {
        solveSplit();
        std::cout << "Total time:  " << metricsTracker.totalTime << "ms" << std::endl;
        if (!conf::trackCholeskySolveStep) metricsTracker.endTracking();
        if (conf::writeMatrix) MatrixParser::writeFullMatrix("./A_chol_result", A);
        return;
    }
    const auto start = clk::now();
    metricsTracker.startTracking();
    hs_handle_t h = gpuQueue.handle();
    const std::size_t bytes = A.matrixData.size() * sizeof(double);
    auto t0 = clk::now();
    gpuQueue.reserve(A.N, A.blockSize);
    if (!A_gpu) {
        void* p = nullptr;
        hs::check(hs_malloc(&p, bytes), "hs_malloc");
        A_gpu = static_cast<double*>(p);
    }
    hs::check(hs_memcpy_htod(h, A_gpu, A.matrixData.data(), bytes), "hs_memcpy_htod");
    metricsTracker.memoryInitTime = ms_since(t0);

    const std::size_t n = static_cast<std::size_t>(A.blockCountXY);
    std::vector<double> potrf(n), trsm(n), update(n);
    int rc;
    if (conf::referenceOrder) {
        // the parity mode: the reference's kernels' arithmetic order, no
        // per-column record (the vectors below hold 0)
        rc = hs_cholesky_ref_order_f64(h, A_gpu, static_cast<int64_t>(A.N), A.blockSize,
                                       conf::referenceOrderFMA ? 1 : 0);
        hs::check(rc, "hs_cholesky_ref_order_f64");
    } else {
        rc = hs_cholesky_f64_trace(h, A_gpu, static_cast<int64_t>(A.N), A.blockSize, potrf.data(), trsm.data(),
                                   update.data());
        hs::check(rc, "hs_cholesky_f64");
    }
    info = rc;
    if (info > 0) std::cerr << "warning: non-positive pivot at row " << info << " (matrix not SPD)" << std::endl;
    // The reference's per-column record (Cholesky.cpp:261-359). SYRK and GEMM
    // are one fused launch here, recorded under matrixMatrixTimes (the
    // diagonal-update vectors hold 0); a launch that applies several panels is
    // recorded at its first column. No rows are shifted or copied between
    // devices at share 1.0: shiftTimes and copyTimes hold 0. The column time
    // is the sum of the column's kernel times (the panel and update streams
    // overlap, so these add up to more than the factorisation's wall time).
    for (std::size_t k = 0; k < n; ++k) {
        metricsTracker.shiftTimes.push_back(0.0);
        metricsTracker.choleskyDiagonalBlockTimes.push_back(potrf[k]);
        metricsTracker.copyTimes.push_back(0.0);
        metricsTracker.triangularSolveTimes_GPU.push_back(trsm[k]);
        metricsTracker.triangularSolveTimes_CPU.push_back(0.0);
        metricsTracker.triangularSolveTimes_total.push_back(trsm[k]);
        metricsTracker.matrixMatrixDiagonalTimes_GPU.push_back(0.0);
        metricsTracker.matrixMatrixDiagonalTimes_CPU.push_back(0.0);
        metricsTracker.matrixMatrixDiagonalTimes_total.push_back(0.0);
        metricsTracker.matrixMatrixTimes_GPU.push_back(update[k]);
        metricsTracker.matrixMatrixTimes_CPU.push_back(0.0);
        metricsTracker.matrixMatrixTimes_total.push_back(update[k]);
        metricsTracker.updateMetrics(k, n, 0, potrf[k] + trsm[k] + update[k], loadBalancer->updateInterval);
    }

    t0 = clk::now();
    hs::check(hs_memcpy_dtoh(h, A.matrixData.data(), A_gpu, bytes), "hs_memcpy_dtoh");
    metricsTracker.resultCopyTime = ms_since(t0);
    metricsTracker.totalTime = ms_since(start);
    std::cout << "Total time:  " << metricsTracker.totalTime << "ms" << std::endl;
    if (!conf::trackCholeskySolveStep) metricsTracker.endTracking();
    if (conf::writeMatrix) MatrixParser::writeFullMatrix("./A_chol_result", A);
}

void Cholesky::writeMetricsToFile() {
    std::string dir = conf::outputPath + "/" + timeString();
    std::filesystem::create_directories(dir);
    metricsTracker.writeJSON(dir);
}

TriangularSystemSolver::TriangularSystemSolver(SymmetricMatrix& A, conf::fp_type* A_gpu, RightHandSide& b,
                                               hs::Queue& cpuQueue, hs::Queue& gpuQueue,
                                               std::shared_ptr<LoadBalancer> loadBalancer)
    : A(A), b(b), A_gpu(A_gpu), cpuQueue(cpuQueue), gpuQueue(gpuQueue), loadBalancer(std::move(loadBalancer)) {}

double TriangularSystemSolver::solve() {
    const auto start = clk::now();
    // the reference solves on the GPU only when the share is exactly 1, on the
    // CPU otherwise (TriangularSystemSolver.cpp:18), with the factor in
    // A.matrixData
    if (conf::initialProportionGPU != 1.0) {
        hs::cpu::tri_solve(A.matrixData.data(), b.rightHandSideData.data(), A.N, A.blockSize);
        const double ms = ms_since(start);
        std::cout << "Time to solve the triangular system: " << ms << "ms" << std::endl;
        if (conf::writeResult) writeResult(".", b.rightHandSideData);
        return ms;
    }
    if (!A_gpu) throw std::invalid_argument("TriangularSystemSolver: no device factor (run Cholesky first)");
    hs_handle_t h = gpuQueue.handle();
    gpuQueue.reserve(A.N, A.blockSize);
    const std::size_t bytes = b.rightHandSideData.size() * sizeof(double);
    hs::DeviceArray<double> db(b.rightHandSideData.size());
    hs::check(hs_memcpy_htod(h, db.data(), b.rightHandSideData.data(), bytes), "hs_memcpy_htod");
    if (conf::referenceOrder)
        hs::check(hs_tri_solve_ref_order_f64(h, A_gpu, db.data(), static_cast<int64_t>(A.N), A.blockSize,
                                             conf::referenceOrderFMA ? 1 : 0),
                  "hs_tri_solve_ref_order_f64");
    else
        hs::check(hs_tri_solve_f64(h, A_gpu, db.data(), static_cast<int64_t>(A.N), A.blockSize), "hs_tri_solve_f64");
    hs::check(hs_memcpy_dtoh(h, b.rightHandSideData.data(), db.data(), bytes), "hs_memcpy_dtoh");
    const double ms = ms_since(start);
    std::cout << "Time to solve the triangular system: " << ms << "ms" << std::endl;
    if (conf::writeResult) writeResult(".", b.rightHandSideData);
    return ms;
}
