// Conjugate gradients, same API as the reference (src/solver/cg/CG.hpp:20-33).
// At share 1.0 the whole iteration runs on the device (hs_cg_f64): packed
// read-once SYMV, fused dot/axpy kernels, one stop-flag read per 8 iterations.
// Below 1.0 the block rows are split between the GPU and the CPU as the
// reference splits them (CG.cpp:43-46).
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "hs/LoadBalancer.hpp"
#include "hs/MetricsTracker.hpp"
#include "hs/Queue.hpp"
#include "hs/RightHandSide.hpp"
#include "hs/SymmetricMatrix.hpp"

class CG {
public:
    CG(SymmetricMatrix& A, RightHandSide& b, hs::Queue& cpuQueue, hs::Queue& gpuQueue,
       std::shared_ptr<LoadBalancer> loadBalancer);

    void solveHeterogeneous();

    SymmetricMatrix& A;
    RightHandSide& b;
    std::vector<conf::fp_type, hs::PinnedAllocator<conf::fp_type>> x;
    hs::Queue& cpuQueue;
    hs::Queue& gpuQueue;
    std::shared_ptr<LoadBalancer> loadBalancer;
    MetricsTracker metricsTracker;

    std::size_t iterations = 0;   // iterations executed by the last solve
    double deltaZero = 0.0;       // r0.r0
    double deltaFinal = 0.0;      // r.r at exit

private:
    void solveSplit();  // GPU share below 1.0: rows split between CPU and GPU
    void finish();      // metrics.json and x_result.txt
};
