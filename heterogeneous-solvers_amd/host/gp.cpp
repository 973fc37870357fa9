// Gaussian-process prediction of the host API: GaussianProcess
// (GaussianProcess.cpp:10-60), MatrixGenerator::generateTestKernelMatrix
// (MatrixGenerator.cpp:163-213, 257-272) and
// MatrixVectorOperations::matrixVectorGP (MatrixVectorOperations.cpp:379-395),
// all computed on the GPU queue through the C ABI.
#include <fstream>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "hs/hs.hpp"

namespace {

// readInputVector (MatrixGenerator.cpp:257-272): up to N values, one per
// line; missing values stay zero, as in the reference.
std::vector<double> readSeries(const std::string& path, std::size_t N) {
    std::vector<double> v(N, 0.0);
    std::ifstream in(path);
    std::string s;
    std::size_t i = 0;
    while (i < N && std::getline(in, s)) v[i++] = std::stod(s);
    return v;
}

// 8-lag window features of a host series on the device
void windowFeatures(hs_handle_t h, const std::vector<double>& t, hs::DeviceArray<double>& X) {
    hs::DeviceArray<double> dt(t.size());
    hs::check(hs_memcpy_htod(h, dt.data(), t.data(), t.size() * sizeof(double)), "hs_memcpy_htod");
    X.resize(t.size() * 8);
    hs::check(hs_window_features_f64(h, dt.data(), static_cast<int64_t>(t.size()), X.data()),
              "hs_window_features_f64");
    hs::check(hs_sync(h), "hs_sync");
}

void requireGPU(const hs::Queue& q, const char* who) {
    if (!q.is_gpu()) {
        throw std::invalid_argument(std::string(who) +
                                    ": this build computes on a GPU queue only (no CPU compute path)");
    }
}

// UtilityFunctions::writeResult (UtilityFunctions.cpp:19-33): one line, values
// separated by spaces, fixed notation with 20 digits
template <typename V>
void writeResultLine(const std::string& dir, const V& x) {
    std::ofstream out(dir + "/x_result.txt");
    out << std::setprecision(20) << std::fixed;
    for (const auto& e : x) out << e << " ";
    out << std::endl;
}

}  // namespace

void MatrixGenerator::generateTestKernelMatrix(std::string& path_train, std::string& path_test, hs::Queue&,
                                               hs::Queue& queueGPU, conf::fp_type* K_star) {
    requireGPU(queueGPU, "generateTestKernelMatrix");
    const std::size_t N = conf::N, M = conf::N_test;
    std::cout << "-- generating Kernel matrix of size " << N << "x" << M << std::endl;
    if (N == 0 || M == 0) throw std::invalid_argument("generateTestKernelMatrix: N and N_test must be positive");
    hs_handle_t h = queueGPU.handle();
    hs::DeviceArray<double> Xtr, Xte;
    windowFeatures(h, readSeries(path_train, N), Xtr);
    windowFeatures(h, readSeries(path_test, M), Xte);
    hs::DeviceArray<double> dK(N * M);
    hs::check(hs_rbf_cross_f64(h, Xtr.data(), static_cast<int64_t>(N), Xte.data(), static_cast<int64_t>(M), 8, 1.0,
                               1.0, dK.data()),
              "hs_rbf_cross_f64");
    hs::check(hs_memcpy_dtoh(h, K_star, dK.data(), N * M * sizeof(double)), "hs_memcpy_dtoh");
}

void MatrixVectorOperations::matrixVectorGP(hs::Queue& queue, conf::fp_type* A, conf::fp_type* b,
                                            conf::fp_type* result, int n, int m) {
    requireGPU(queue, "matrixVectorGP");
    if (n <= 0 || m <= 0) throw std::invalid_argument("matrixVectorGP: n and m must be positive");
    hs_handle_t h = queue.handle();
    const std::size_t nn = static_cast<std::size_t>(n), mm = static_cast<std::size_t>(m);
    hs::DeviceArray<double> dA(nn * mm), dx(nn), dy(mm);
    hs::check(hs_memcpy_htod(h, dA.data(), A, nn * mm * sizeof(double)), "hs_memcpy_htod");
    hs::check(hs_memcpy_htod(h, dx.data(), b, nn * sizeof(double)), "hs_memcpy_htod");
    hs::check(hs_gemv_rows_f64(h, dA.data(), dx.data(), dy.data(), n, m), "hs_gemv_rows_f64");
    hs::check(hs_memcpy_dtoh(h, result, dy.data(), mm * sizeof(double)), "hs_memcpy_dtoh");
}

GaussianProcess::GaussianProcess(SymmetricMatrix& A, RightHandSide& train_y, std::string& path_train,
                                 std::string& path_test, hs::Queue& cpuQueue, hs::Queue& gpuQueue,
                                 std::shared_ptr<LoadBalancer> loadBalancer)
    : A(A),
      train_y(train_y),
      path_train(path_train),
      path_test(path_test),
      cpuQueue(cpuQueue),
      gpuQueue(gpuQueue),
      loadBalancer(std::move(loadBalancer)) {}

void GaussianProcess::start() {
    requireGPU(gpuQueue, "GaussianProcess");
    if (conf::algorithm == "cg") {
        CG cg(A, train_y, cpuQueue, gpuQueue, loadBalancer);
        cg.solveHeterogeneous();
        train_y.rightHandSideData.assign(cg.x.begin(), cg.x.end());
    } else if (conf::algorithm == "cholesky") {
        Cholesky cholesky(A, cpuQueue, gpuQueue, loadBalancer);
        cholesky.solve_heterogeneous();
        TriangularSystemSolver solver(A, cholesky.A_gpu, train_y, cpuQueue, gpuQueue, loadBalancer);
        const double solveTime = solver.solve();
        cholesky.metricsTracker.solveTime = solveTime;
        cholesky.writeMetricsToFile();
    } else {
        throw std::runtime_error("Invalid algorithm: " + conf::algorithm);
    }

    const std::size_t N = conf::N, M = conf::N_test;
    if (M == 0) throw std::invalid_argument("GaussianProcess: conf::N_test (--test_size) must be positive");
    std::cout << "-- predicting at " << M << " test inputs" << std::endl;
    hs_handle_t h = gpuQueue.handle();
    gpuQueue.reserve(N, conf::matrixBlockSize, M);
    hs::DeviceArray<double> Xtr, Xte;
    windowFeatures(h, readSeries(path_train, N), Xtr);
    windowFeatures(h, readSeries(path_test, M), Xte);
    hs::DeviceArray<double> dalpha(N), dmean(M);
    hs::check(hs_memcpy_htod(h, dalpha.data(), train_y.rightHandSideData.data(), N * sizeof(double)),
              "hs_memcpy_htod");
    hs::check(hs_gp_predict_f64(h, Xtr.data(), static_cast<int64_t>(N), Xte.data(), static_cast<int64_t>(M), 8, 1.0,
                                1.0, dalpha.data(), dmean.data()),
              "hs_gp_predict_f64");
    result.assign(M, 0.0);
    hs::check(hs_memcpy_dtoh(h, result.data(), dmean.data(), M * sizeof(double)), "hs_memcpy_dtoh");
    if (conf::writeResult) writeResultLine(".", result);
}
