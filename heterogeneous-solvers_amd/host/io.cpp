// Text formats (MatrixParser.cpp:11-274) and the GP kernel-matrix builder
// (MatrixGenerator.cpp:73-162, 215-240) of the host API.
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "hs/hs.hpp"

std::vector<conf::fp_type> MatrixParser::getRowValuesFromString(const std::string& rowString) {
    std::vector<conf::fp_type> values;
    std::size_t last = 0;
    auto parse = [&](const std::string& s) {
        try {
            std::size_t used = 0;
            const double v = std::stod(s, &used);
            values.push_back(v);
        } catch (...) {
            throw std::invalid_argument("Error while parsing string '" + s + "'");
        }
    };
    for (std::size_t i = 0; i < rowString.size(); ++i) {
        if (rowString[i] == ';') {
            parse(rowString.substr(last, i - last));
            last = i + 1;
        }
    }
    parse(rowString.substr(last));
    return values;
}

SymmetricMatrix MatrixParser::parseSymmetricMatrix(std::string& path, hs::Queue& queue) {
    std::ifstream in(path);
    std::string row;
    std::getline(in, row);
    if (row.empty() || row[0] != '#') throw std::invalid_argument("Invalid matrix format. First line has to be '# <N>!'");
    const std::size_t N = std::stoul(row.substr(2));
    conf::N = N;
    SymmetricMatrix A(N, static_cast<int>(conf::matrixBlockSize), queue);
    const int b = A.blockSize;
    std::size_t i = 0;
    while (i < N && std::getline(in, row)) {
        const auto vals = getRowValuesFromString(row);
        if (vals.size() != i + 1) throw std::invalid_argument("row " + std::to_string(i) + " has the wrong length");
        const int I = static_cast<int>(i / b), il = static_cast<int>(i % b);
        for (std::size_t j = 0; j <= i; ++j) {
            const int J = static_cast<int>(j / b), jl = static_cast<int>(j % b);
            const std::size_t base = A.blockID(I, J) * b * b;
            A.matrixData[base + static_cast<std::size_t>(il) * b + jl] = vals[j];
            if (I == J) A.matrixData[base + static_cast<std::size_t>(jl) * b + il] = vals[j];
        }
        ++i;
    }
    return A;
}

RightHandSide MatrixParser::parseRightHandSide(std::string& path, hs::Queue& queue) {
    std::ifstream in(path);
    std::string row;
    std::getline(in, row);
    if (row.empty() || row[0] != '#')
        throw std::invalid_argument("Invalid right-hand side format. First line has to be '# <N>!'");
    const std::size_t N = std::stoul(row.substr(2));
    std::getline(in, row);
    const auto vals = getRowValuesFromString(row);
    if (vals.size() < N) throw std::invalid_argument("right-hand side has fewer than N values");
    RightHandSide b(N, static_cast<int>(conf::matrixBlockSize), queue);
    for (std::size_t i = 0; i < N; ++i) b.rightHandSideData[i] = vals[i];
    return b;
}

// One value as the reference writes it: fixed notation, a leading space for
// values >= 0 (so columns line up with negative ones), ';' after each.
static void putValue(std::ostream& out, double v) {
    if (v >= 0) out << ' ';
    out << v << ';';
}

// All n b rows and columns (padding included), fixed with 20 digits: row i
// takes blocks J <= I row-wise, then blocks J > I from the stored (J, I)
// transposed (MatrixParser.cpp:214-274). The file is `path` itself.
void MatrixParser::writeFullMatrix(const std::string& path, const SymmetricMatrix& A) {
    std::ofstream out(path);
    out << std::setprecision(20) << std::fixed;
    const int b = A.blockSize, n = A.blockCountXY;
    const std::size_t bb = static_cast<std::size_t>(b) * b;
    for (int row = 0; row < n * b; ++row) {
        const int I = row / b, il = row % b;
        for (int J = 0; J <= I; ++J) {
            const double* blk = A.matrixData.data() + A.blockID(I, J) * bb + static_cast<std::size_t>(il) * b;
            for (int j = 0; j < b; ++j) putValue(out, blk[j]);
        }
        for (int J = I + 1; J < n; ++J) {
            const double* blk = A.matrixData.data() + A.blockID(J, I) * bb;
            for (int j = 0; j < b; ++j) putValue(out, blk[static_cast<std::size_t>(j) * b + il]);
        }
        out << std::endl;
    }
}

// The stored blocks of each row's block row, fixed with 10 digits, a tab
// after each block and an empty line before each block row
// (MatrixParser.cpp:170-211).
void MatrixParser::writeBlockedMatrix(const std::string& path, const SymmetricMatrix& A) {
    std::ofstream out(path);
    out << std::setprecision(10) << std::fixed;
    const int b = A.blockSize, n = A.blockCountXY;
    const std::size_t bb = static_cast<std::size_t>(b) * b;
    for (int row = 0; row < n * b; ++row) {
        if (row % b == 0) out << std::endl;
        const int I = row / b, il = row % b;
        for (int J = 0; J <= I; ++J) {
            const double* blk = A.matrixData.data() + A.blockID(I, J) * bb + static_cast<std::size_t>(il) * b;
            for (int j = 0; j < b; ++j) putValue(out, blk[j]);
            out << '\t';
        }
        out << std::endl;
    }
}

static std::vector<double> readColumn(const std::string& path, std::size_t N) {
    std::ifstream in(path);
    std::vector<double> v;
    v.reserve(N);
    std::string s;
    while (v.size() < N && std::getline(in, s)) v.push_back(std::stod(s));
    if (v.size() != N) throw std::runtime_error("Not enough data available!");
    return v;
}

SymmetricMatrix MatrixGenerator::generateSPDMatrix(std::string& path, hs::Queue&, hs::Queue& queueGPU) {
    const std::size_t N = conf::N;
    std::cout << "-- generating SPD matrix of size " << N << "x" << N << std::endl;
    const std::vector<double> t = readColumn(path, N);
    SymmetricMatrix A(N, static_cast<int>(conf::matrixBlockSize), queueGPU);
    hs::DeviceArray<double> dt(N), dA(A.matrixData.size());
    hs_handle_t h = queueGPU.handle();
    queueGPU.reserve(N, A.blockSize);
    hs::check(hs_memcpy_htod(h, dt.data(), t.data(), N * sizeof(double)), "hs_memcpy_htod");
    hs::check(hs_rbf_window_f64(h, dt.data(), static_cast<int64_t>(N), A.blockSize, dA.data()), "hs_rbf_window_f64");
    hs::check(hs_memcpy_dtoh(h, A.matrixData.data(), dA.data(), A.matrixData.size() * sizeof(double)),
              "hs_memcpy_dtoh");
    return A;
}

SymmetricMatrix MatrixGenerator::generateSPDMatrix(const double* X, std::size_t N, int d, double lengthscale,
                                                   double signalVariance, double noiseVariance, int blockSize,
                                                   hs::Queue& queueGPU) {
    SymmetricMatrix A(N, blockSize, queueGPU);
    hs::DeviceArray<double> dX(N * d), dA(A.matrixData.size());
    hs_handle_t h = queueGPU.handle();
    hs::check(hs_memcpy_htod(h, dX.data(), X, N * d * sizeof(double)), "hs_memcpy_htod");
    hs::check(hs_rbf_assemble_f64(h, dX.data(), static_cast<int64_t>(N), d, blockSize, lengthscale, signalVariance,
                                  noiseVariance, dA.data()),
              "hs_rbf_assemble_f64");
    hs::check(hs_memcpy_dtoh(h, A.matrixData.data(), dA.data(), A.matrixData.size() * sizeof(double)),
              "hs_memcpy_dtoh");
    return A;
}

RightHandSide MatrixGenerator::parseRHS_GP(std::string& path, hs::Queue& queue) {
    const std::size_t N = conf::N;
    const std::vector<double> v = readColumn(path, N);
    RightHandSide b(N, static_cast<int>(conf::matrixBlockSize), queue);
    for (std::size_t i = 0; i < N; ++i) b.rightHandSideData[i] = v[i];
    return b;
}
