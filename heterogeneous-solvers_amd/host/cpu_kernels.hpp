// The CPU half of a heterogeneous solve (GPU share below 1.0): FP64 kernels on
// the host copy of the packed lower-block-triangle matrix, OpenMP over blocks
// and rows. They take the reference's block coordinates (the CPU queue's
// kernels of MatrixOperations.cpp:7-62, MatrixMatrixOperations.cpp:109-144,
// 307-361, 845-906, MatrixVectorOperations.cpp:194-377 and
// VectorOperations.cpp:45-82, 187-238) and compute the same quantities in
// their own loop order: every inner loop is a dot product over two
// contiguous rows of a row-major block.
#pragma once

#include <cstddef>
#include <cstdint>

namespace hs {
namespace cpu {

// packed ID of block (I, J), I >= J, among n block columns
inline std::size_t blockID(int I, int J, int n) {
    return static_cast<std::size_t>(J) * n - static_cast<std::size_t>(J) * (J - 1) / 2 + (I - J);
}

// Cholesky of the diagonal block of block row k (stored at block `id`), in
// place: lower factor, upper triangle zeroed; rows >= N (padding) untouched.
// Returns 0, or the 1-based global row of the first non-positive pivot.
int64_t potrf(double* A, std::size_t N, int b, std::size_t id, int k);

// B <- B L_kk^{-T} for the `count` panel blocks (rows start .. start+count-1)
// of block column k, whose diagonal block is `id`.
void trsm(double* A, int b, std::size_t id, int k, int start, int count);

// D_I <- D_I - L_Ik L_Ik^T (lower triangle) for I in [start, start + count).
void syrk(double* A, int b, int k, int start, int count, int n);

// A_IJ <- A_IJ - L_Ik L_Jk^T for I in [start, start + count), k < J < I.
void gemm(double* A, int b, int k, int start, int count, int n);

// y_I = sum_J A_IJ x_J for block rows I in [start, start + count), all J.
void symv_rows(const double* A, const double* x, double* y, int b, int start, int count, int n);

// sum x_i y_i over [begin, end): per-thread partials of static chunks,
// added in thread order.
double dot(const double* x, const double* y, std::size_t begin, std::size_t end);

// r = x + a y and r = x - y over [begin, end)
void scale_add(const double* x, double a, const double* y, double* r, std::size_t begin, std::size_t end);
void sub(const double* x, const double* y, double* r, std::size_t begin, std::size_t end);

// L L^T x = y in place on x (n b entries; rows >= N stay as they are)
void tri_solve(const double* L, double* x, std::size_t N, int b);

}  // namespace cpu
}  // namespace hs
