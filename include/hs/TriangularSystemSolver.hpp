// Forward and backward blocked substitution with the factor on the device,
// same API as the reference (TriangularSystemSolver.hpp:17-33): the solution
// overwrites b, solve() returns the elapsed milliseconds.
#pragma once

#include <memory>

#include "hs/LoadBalancer.hpp"
#include "hs/Queue.hpp"
#include "hs/RightHandSide.hpp"
#include "hs/SymmetricMatrix.hpp"

class TriangularSystemSolver {
public:
    TriangularSystemSolver(SymmetricMatrix& A, conf::fp_type* A_gpu, RightHandSide& b, hs::Queue& cpuQueue,
                           hs::Queue& gpuQueue, std::shared_ptr<LoadBalancer> loadBalancer);
    double solve();

    SymmetricMatrix& A;
    RightHandSide& b;
    conf::fp_type* A_gpu;
    hs::Queue& cpuQueue;
    hs::Queue& gpuQueue;
    std::shared_ptr<LoadBalancer> loadBalancer;
};
