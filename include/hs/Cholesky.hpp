// Right-looking blocked Cholesky, same API as the reference
// (src/solver/cholesky/Cholesky.hpp:17-35). solve_heterogeneous() factors
// A = L L^T on the device (hs_cholesky_f64: FP64 MFMA trailing update with a
// one-panel look-ahead), keeps L on the device in A_gpu for the triangular
// solves and copies it back into A.matrixData, as the reference does at
// share 1.0 (Cholesky.cpp:389-395). Below share 1.0 the trailing block rows of
// every column are split between the CPU and the GPU (Cholesky.cpp:56-259),
// and L ends up in A.matrixData.
#pragma once

#include <cstdint>
#include <memory>

#include "hs/LoadBalancer.hpp"
#include "hs/MetricsTracker.hpp"
#include "hs/Queue.hpp"
#include "hs/SymmetricMatrix.hpp"

class Cholesky {
public:
    Cholesky(SymmetricMatrix& A, hs::Queue& cpuQueue, hs::Queue& gpuQueue, std::shared_ptr<LoadBalancer> loadBalancer);
    ~Cholesky();

    void solve_heterogeneous();
    void writeMetricsToFile();

    SymmetricMatrix& A;
    conf::fp_type* A_gpu = nullptr;
    hs::Queue& cpuQueue;
    hs::Queue& gpuQueue;
    std::shared_ptr<LoadBalancer> loadBalancer;
    MetricsTracker metricsTracker;

    // 0, or the 1-based row of the first non-positive pivot (the reference
    // continues with NaN there; this build reports it)
    int64_t info = 0;

private:
    void solveSplit();  // GPU share below 1.0: trailing rows split between CPU and GPU
};
