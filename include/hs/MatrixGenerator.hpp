// GP RBF kernel-matrix builder (src/utility/MatrixGenerator.hpp:30, 50).
// generateSPDMatrix reads the reference's 1-D input file (8-lag window
// features, sigma_f^2 = 1, l = 1, noise 0.01) and assembles the packed matrix
// on the GPU; the overload takes explicit features and hyperparameters.
#pragma once

#include <cstddef>
#include <string>

#include "hs/Queue.hpp"
#include "hs/RightHandSide.hpp"
#include "hs/SymmetricMatrix.hpp"

class MatrixGenerator {
public:
    static SymmetricMatrix generateSPDMatrix(std::string& path, hs::Queue& queue, hs::Queue& queueGPU);
    static SymmetricMatrix generateSPDMatrix(const double* X, std::size_t N, int d, double lengthscale,
                                             double signalVariance, double noiseVariance, int blockSize,
                                             hs::Queue& queueGPU);
    static RightHandSide parseRHS_GP(std::string& path, hs::Queue& queue);
    // K*[j * conf::N + i] for the conf::N training and conf::N_test test inputs
    // (8-lag windows of the two series, sigma_f^2 = 1, l = 1) into the host
    // array K_star, computed on queueGPU (MatrixGenerator.cpp:163-213).
    static void generateTestKernelMatrix(std::string& path_train, std::string& path_test, hs::Queue& queue,
                                         hs::Queue& queueGPU, conf::fp_type* K_star);
};
