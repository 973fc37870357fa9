// CPU kernels of the heterogeneous solves (declarations and conventions in
// cpu_kernels.hpp).
#include "cpu_kernels.hpp"

#include <cmath>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace hs {
namespace cpu {

namespace {

inline double rowdot(const double* a, const double* c, int len) {
    double s = 0.0;
    for (int t = 0; t < len; ++t) s += a[t] * c[t];
    return s;
}

int threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

}  // namespace

// Row-oriented (Crout) factorisation: column j of L from the rows above it,
// each entry one dot product over the already factored part of two rows.
int64_t potrf(double* A, std::size_t N, int b, std::size_t id, int k) {
    double* D = A + id * b * b;
    const std::size_t row0 = static_cast<std::size_t>(k) * b;
    const int m = static_cast<int>(std::min<std::size_t>(b, N > row0 ? N - row0 : 0));  // non-padding rows
    int64_t info = 0;
    for (int j = 0; j < m; ++j) {
        double* Lj = D + static_cast<std::size_t>(j) * b;
        const double djj = Lj[j] - rowdot(Lj, Lj, j);
        if (!(djj > 0.0) && info == 0) info = static_cast<int64_t>(row0) + j + 1;
        const double ljj = std::sqrt(djj);
        Lj[j] = ljj;
#pragma omp parallel for schedule(static) if (m - j > 64)
        for (int i = j + 1; i < m; ++i) {
            double* Li = D + static_cast<std::size_t>(i) * b;
            Li[j] = (Li[j] - rowdot(Li, Lj, j)) / ljj;
        }
    }
    for (int i = 0; i < m; ++i)
        for (int j = i + 1; j < b; ++j) D[static_cast<std::size_t>(i) * b + j] = 0.0;
    return info;
}

// x L^T = row for every row of every panel block: x_j = (row_j - x_[0,j) .
// L_j,[0,j)) / L_jj, in place.
void trsm(double* A, int b, std::size_t id, int k, int start, int count) {
    const double* L = A + id * b * b;
    std::vector<double> rdiag(b);
    for (int j = 0; j < b; ++j) rdiag[j] = L[static_cast<std::size_t>(j) * b + j];
    const std::size_t first = id + (start - k);
#pragma omp parallel for collapse(2) schedule(static)
    for (int g = 0; g < count; ++g)
        for (int r = 0; r < b; ++r) {
            double* row = A + (first + g) * b * b + static_cast<std::size_t>(r) * b;
            for (int j = 0; j < b; ++j) {
                const double* Lj = L + static_cast<std::size_t>(j) * b;
                row[j] = (row[j] - rowdot(row, Lj, j)) / rdiag[j];
            }
        }
}

void syrk(double* A, int b, int k, int start, int count, int n) {
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int g = 0; g < count; ++g)
        for (int i = 0; i < b; ++i) {
            const int I = start + g;
            double* D = A + blockID(I, I, n) * b * b;
            const double* C = A + blockID(I, k, n) * b * b;
            const double* Ci = C + static_cast<std::size_t>(i) * b;
            for (int j = 0; j <= i; ++j) D[static_cast<std::size_t>(i) * b + j] -= rowdot(Ci, C + static_cast<std::size_t>(j) * b, b);
        }
}

void gemm(double* A, int b, int k, int start, int count, int n) {
    // the (I, J) pairs of the update, flattened for the parallel loop
    std::vector<int> PI, PJ;
    for (int I = start; I < start + count; ++I)
        for (int J = k + 1; J < I; ++J) {
            PI.push_back(I);
            PJ.push_back(J);
        }
    const int pairs = static_cast<int>(PI.size());
#pragma omp parallel for collapse(2) schedule(dynamic, 8)
    for (int p = 0; p < pairs; ++p)
        for (int i = 0; i < b; ++i) {
            const int I = PI[p], J = PJ[p];
            double* Crow = A + blockID(I, J, n) * b * b + static_cast<std::size_t>(i) * b;
            const double* Li = A + blockID(I, k, n) * b * b + static_cast<std::size_t>(i) * b;
            const double* LJ = A + blockID(J, k, n) * b * b;
            for (int j = 0; j < b; ++j) Crow[j] -= rowdot(Li, LJ + static_cast<std::size_t>(j) * b, b);
        }
}

// Each block row in one task: blocks J <= I row by row (a dot per row), blocks
// J > I as A_JI^T x_J, added row of A_JI by row (an axpy per row), so that
// every block is streamed in storage order.
void symv_rows(const double* A, const double* x, double* y, int b, int start, int count, int n) {
#pragma omp parallel for schedule(dynamic)
    for (int I = start; I < start + count; ++I) {
        std::vector<double> acc(b, 0.0);
        for (int J = 0; J < n; ++J) {
            const double* xJ = x + static_cast<std::size_t>(J) * b;
            if (J <= I) {
                const double* blk = A + blockID(I, J, n) * b * b;
                for (int i = 0; i < b; ++i) acc[i] += rowdot(blk + static_cast<std::size_t>(i) * b, xJ, b);
            } else {
                const double* blk = A + blockID(J, I, n) * b * b;
                for (int j = 0; j < b; ++j) {
                    const double xj = xJ[j];
                    const double* row = blk + static_cast<std::size_t>(j) * b;
                    for (int i = 0; i < b; ++i) acc[i] += row[i] * xj;
                }
            }
        }
        double* yI = y + static_cast<std::size_t>(I) * b;
        for (int i = 0; i < b; ++i) yI[i] = acc[i];
    }
}

double dot(const double* x, const double* y, std::size_t begin, std::size_t end) {
    const int T = threads();
    std::vector<double> part(T, 0.0);
#pragma omp parallel num_threads(T)
    {
        const int t = thread_id();
        const std::size_t len = end > begin ? end - begin : 0;
        const std::size_t lo = begin + len * t / T, hi = begin + len * (t + 1) / T;
        double s = 0.0;
        for (std::size_t i = lo; i < hi; ++i) s += x[i] * y[i];
        part[t] = s;
    }
    double s = 0.0;
    for (double v : part) s += v;
    return s;
}

void scale_add(const double* x, double a, const double* y, double* r, std::size_t begin, std::size_t end) {
#pragma omp parallel for schedule(static)
    for (std::size_t i = begin; i < end; ++i) r[i] = x[i] + y[i] * a;
}

void sub(const double* x, const double* y, double* r, std::size_t begin, std::size_t end) {
#pragma omp parallel for schedule(static)
    for (std::size_t i = begin; i < end; ++i) r[i] = x[i] - y[i];
}

// Forward: z_J = L_JJ^{-1} y_J, then y_I -= L_IJ z_J below it; backward:
// x_J = L_JJ^{-T} z_J, then z_I -= L_JI^T x_J above it. The block-row
// updates run in parallel.
void tri_solve(const double* L, double* x, std::size_t N, int b) {
    const int n = static_cast<int>((N + b - 1) / b);
    auto rows = [&](int J) {
        const std::size_t r0 = static_cast<std::size_t>(J) * b;
        return static_cast<int>(std::min<std::size_t>(b, N > r0 ? N - r0 : 0));
    };
    for (int J = 0; J < n; ++J) {
        const double* D = L + blockID(J, J, n) * b * b;
        double* xJ = x + static_cast<std::size_t>(J) * b;
        const int m = rows(J);
        for (int i = 0; i < m; ++i) {
            const double* Di = D + static_cast<std::size_t>(i) * b;
            xJ[i] = (xJ[i] - rowdot(Di, xJ, i)) / Di[i];
        }
#pragma omp parallel for schedule(static)
        for (int I = J + 1; I < n; ++I) {
            const double* blk = L + blockID(I, J, n) * b * b;
            double* xI = x + static_cast<std::size_t>(I) * b;
            for (int i = 0; i < b; ++i) xI[i] -= rowdot(blk + static_cast<std::size_t>(i) * b, xJ, b);
        }
    }
    for (int J = n - 1; J >= 0; --J) {
        const double* D = L + blockID(J, J, n) * b * b;
        double* xJ = x + static_cast<std::size_t>(J) * b;
        const int m = rows(J);
        for (int i = m - 1; i >= 0; --i) {
            double s = xJ[i];
            for (int t = i + 1; t < m; ++t) s -= D[static_cast<std::size_t>(t) * b + i] * xJ[t];
            xJ[i] = s / D[static_cast<std::size_t>(i) * b + i];
        }
#pragma omp parallel for schedule(static)
        for (int I = 0; I < J; ++I) {
            const double* blk = L + blockID(J, I, n) * b * b;  // L_JI: x_I -= L_JI^T x_J
            double* xI = x + static_cast<std::size_t>(I) * b;
            for (int j = 0; j < b; ++j) {
                const double xj = xJ[j];
                const double* row = blk + static_cast<std::size_t>(j) * b;
                for (int i = 0; i < b; ++i) xI[i] -= row[i] * xj;
            }
        }
    }
}

}  // namespace cpu
}  // namespace hs
