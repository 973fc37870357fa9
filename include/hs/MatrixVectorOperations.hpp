// The GP mat-vec of the reference's MatrixVectorOperations
// (src/kernel/MatrixVectorOperations.hpp, .cpp:379-395): result = A b for a
// row-major m x n host matrix A. It runs on a GPU queue (hs_gemv_rows_f64);
// a CPU queue is rejected with std::invalid_argument (this build has no CPU
// compute path).
#pragma once

#include "hs/Configuration.hpp"
#include "hs/Queue.hpp"

class MatrixVectorOperations {
public:
    static void matrixVectorGP(hs::Queue& queue, conf::fp_type* A, conf::fp_type* b, conf::fp_type* result, int n,
                               int m);
};
