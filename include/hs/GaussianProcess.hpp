// Gaussian-process prediction pipeline (src/gaussianProcess/GaussianProcess.hpp:
// 12-37): solve K alpha = y with CG or Cholesky, then the predictive mean
// K* alpha at the test inputs. Same constructor and start(); the prediction
// runs on the GPU queue (hs_gp_predict_f64) without storing K*, and the mean
// stays available in `result` after start() (the reference keeps it local and
// writes it with conf::writeResult).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "hs/LoadBalancer.hpp"
#include "hs/Queue.hpp"
#include "hs/RightHandSide.hpp"
#include "hs/SymmetricMatrix.hpp"

class GaussianProcess {
public:
    GaussianProcess(SymmetricMatrix& A, RightHandSide& train_y, std::string& path_train, std::string& path_test,
                    hs::Queue& cpuQueue, hs::Queue& gpuQueue, std::shared_ptr<LoadBalancer> loadBalancer);

    SymmetricMatrix& A;        // training-training kernel
    RightHandSide& train_y;    // training targets; holds alpha = K^{-1} y after start()
    std::string& path_train;   // training series (one value per line)
    std::string& path_test;    // test series (one value per line)
    hs::Queue& cpuQueue;
    hs::Queue& gpuQueue;
    std::shared_ptr<LoadBalancer> loadBalancer;

    std::vector<double> result;  // predictive mean at the conf::N_test test inputs

    // CG or Cholesky + triangular solves by conf::algorithm, then the
    // prediction (GaussianProcess.cpp:20-60).
    void start();
};
